"""CPU model of the merge-path control flow of the SpMM kernels, step for step
as they walk a chunk: k_spmm_stream (one nonzero per step) and k_spmm_slice (a
pair of nonzeros per step, lower and upper half-wave) — full batches with one
row-end test per batch, then the tail — and k_spmm_narrow (each lane walks a
run of merge items; a segmented scan joins rows cut between lanes).  Each chunk
must write exactly the rows it completes, read only the row ends it staged,
and leave the partial sum of its last row to the carry; the sum of written rows
and carries must equal A·1 exactly (small-integer values)."""
import numpy as np
import pytest

from armsparse_amd import data

INF = 1 << 62


def merge_search(rp, d, M, nnz):
    lo, hi = max(0, d - nnz), min(d, M)
    while lo < hi:
        mid = (lo + hi) // 2
        if rp[mid + 1] <= d - mid - 1:
            lo = mid + 1
        else:
            hi = mid
    return lo


def run_model(rp, vals, chunk, U, pairs, last_bound="fixed"):
    M, nnz = len(rp) - 1, int(rp[-1])
    out = np.zeros(M, np.int64)
    emitted = np.zeros(M, np.int64)
    carries = {}
    lanes = 2 if pairs else 1
    step_len = 2 * U if pairs else U
    for d0 in range(0, M + nnz, chunk):
        d1 = min(d0 + chunk, M + nnz)
        i0, i1 = merge_search(rp, d0, M, nnz), merge_search(rp, d1, M, nnz)
        j0, j1 = d0 - i0, d1 - i1
        nz, nr = j1 - j0, i1 - i0 + 1
        s_re = [int(rp[i0 + k + 1]) - j0 if i0 + k < M else INF for k in range(nr)]
        st = dict(rr=0, re=s_re[0], acc=[0] * lanes)

        def emit():
            assert st["rr"] < nr - 1, "emitted the chunk's carry row or past the staged rows"
            r = i0 + st["rr"]
            out[r] += sum(st["acc"])
            emitted[r] += 1
            st["acc"] = [0] * lanes
            st["rr"] += 1
            st["re"] = s_re[st["rr"]] if st["rr"] < nr else INF

        def step(k2, last, has):
            done = [not h for h in has]
            while st["re"] <= last:
                for h in range(lanes):
                    if not done[h] and k2 + h < st["re"]:
                        st["acc"][h] += int(vals[j0 + k2 + h])
                        done[h] = True
                emit()
            for h in range(lanes):
                if not done[h]:
                    st["acc"][h] += int(vals[j0 + k2 + h])

        kend = (nz // step_len) * step_len
        for kb in range(0, kend, step_len):
            if st["re"] >= kb + step_len:  # fast path: no row ends inside the batch
                for k in range(kb, kb + step_len):
                    st["acc"][(k - kb) % lanes] += int(vals[j0 + k])
            else:
                for k2 in range(kb, kb + step_len, lanes):
                    step(k2, k2 + lanes - 1, [True] * lanes)
        for k2 in range(kend, nz, lanes):
            has = [k2 + h < nz for h in range(lanes)]
            if last_bound == "fixed":
                last = max(k2 + h for h in range(lanes) if has[h])
            else:  # the bound the slice kernel first had: the pair's upper index even when it has no nonzero
                last = k2 + lanes - 1
            step(k2, last, has)
        while st["rr"] < nr - 1:
            emit()
        if i1 < M and j1 > max(int(rp[i1]), j0):
            carries.setdefault(i1, []).append(sum(st["acc"]))
    for r, cs in carries.items():
        out[r] += sum(cs)
    return out, emitted


@pytest.fixture(scope="module")
def power_law():
    a = data.rmat_csr(11, 1 << 14, seed=1234)
    vals = np.random.default_rng(0).integers(-9, 10, a.nnz)
    return a.rowptr.astype(np.int64), vals


@pytest.mark.parametrize("pairs", [False, True], ids=["stream", "slice"])
@pytest.mark.parametrize("chunk,U", [(64, 4), (64, 8), (128, 16), (512, 8), (2048, 16)])
def test_merge_control_flow(power_law, pairs, chunk, U):
    rp, vals = power_law
    out, emitted = run_model(rp, vals, chunk, U, pairs)
    expect = np.add.reduceat(np.append(vals, 0), rp[:-1])
    expect[np.diff(rp) == 0] = 0
    assert np.array_equal(out, expect)
    assert np.all(emitted == 1)


def test_model_catches_the_unbounded_tail(power_law):
    """With the pair's upper index as the bound even when the upper half holds no
    nonzero, a chunk whose carry row ends exactly at the chunk end emits it and
    reads the row end past the staged ones."""
    rp, vals = power_law
    with pytest.raises(AssertionError, match="carry row or past"):
        run_model(rp, vals, 64, 4, True, last_bound="old")


def run_narrow_model(rp, vals, ipt):
    """k_spmm_narrow: each of 64 lanes walks ipt items of the wave's merge tile;
    rows that start and end in a lane are written by it, a lane's first finished
    row also gets the partials of the lanes before it (segmented scan), and the
    tile's last open row goes to the carry."""
    M, nnz = len(rp) - 1, int(rp[-1])
    tile = 64 * ipt
    out = np.zeros(M, np.int64)
    written = np.zeros(M, np.int64)
    carries = {}
    for d0 in range(0, M + nnz, tile):
        d1 = min(d0 + tile, M + nnz)
        i0, i1 = merge_search(rp, d0, M, nnz), merge_search(rp, d1, M, nnz)
        j0, j1 = d0 - i0, d1 - i1
        nz, nrows = j1 - j0, i1 - i0
        re = [int(rp[i0 + k + 1]) - j0 for k in range(nrows)]
        items = nrows + nz
        keys, accs, firsts, starts, ends = [], [], [], [], []
        for lane in range(64):
            ld = min(lane * ipt, items)
            lo, hi = max(0, ld - nz), min(ld, nrows)
            while lo < hi:
                mid = (lo + hi) // 2
                if re[mid] <= ld - mid - 1:
                    lo = mid + 1
                else:
                    hi = mid
            ri, rj = lo, ld - lo
            ri_start, acc, first = ri, 0, 0
            for _ in range(ipt):
                if ri + rj < items:
                    if ri < nrows and re[ri] <= rj:
                        if ri == ri_start:
                            first = acc
                        else:
                            out[i0 + ri] += acc
                            written[i0 + ri] += 1
                        acc = 0
                        ri += 1
                    else:
                        acc += int(vals[j0 + rj])
                        rj += 1
            keys.append(ri); accs.append(acc); firsts.append(first); starts.append(ri_start); ends.append(ri)
        S = list(accs)  # segmented inclusive scan by key
        for lane in range(1, 64):
            if keys[lane - 1] == keys[lane]:
                S[lane] += S[lane - 1]
        for lane in range(64):
            if ends[lane] > starts[lane]:
                f = firsts[lane] + (S[lane - 1] if lane > 0 and keys[lane - 1] == starts[lane] else 0)
                out[i0 + starts[lane]] += f
                written[i0 + starts[lane]] += 1
        if i1 < M and j1 > max(int(rp[i1]), j0):
            assert keys[63] == nrows
            carries.setdefault(i1, []).append(S[63])
    for r, cs in carries.items():
        out[r] += sum(cs)
    return out, written


@pytest.mark.parametrize("ipt", [4, 8, 16])
def test_narrow_control_flow(power_law, ipt):
    rp, vals = power_law
    out, written = run_narrow_model(rp, vals, ipt)
    expect = np.add.reduceat(np.append(vals, 0), rp[:-1])
    expect[np.diff(rp) == 0] = 0
    assert np.array_equal(out, expect)
    assert np.all(written == 1)


def scatter_partition(rp, chunk):
    """k_merge_partition: row i (and the extra i = M) writes the boundaries w
    with e(i-1) <= w·chunk < e(i), e(i) = rowptr[i+1] + i + 1."""
    M, nnz = len(rp) - 1, int(rp[-1])
    nwaves = -(-(M + nnz) // chunk)
    part = np.full(nwaves + 1, -1, np.int64)
    for i in range(M + 1):
        lo = 0 if i == 0 else int(rp[i]) + i
        hi = int(rp[i + 1]) + i + 1 if i < M else nwaves * chunk + 1
        wb, we = -(-lo // chunk), min(-(-hi // chunk), nwaves + 1)
        part[wb:we] = i
    return part, nwaves


def _matrices(power_law):
    rp, _ = power_law
    yield "rmat", rp
    yield "empty_rows_at_ends", np.array([0, 0, 0, 5, 5, 9, 64, 64, 64], np.int64)
    yield "one_dense_row", np.array([0, 4096], np.int64)
    yield "no_nonzeros", np.zeros(300, np.int64)
    lens = np.random.default_rng(5).integers(0, 3, 5000)  # many empty and short rows
    yield "short_rows", np.concatenate([[0], np.cumsum(lens)])


@pytest.mark.parametrize("chunk", [64, 256, 512, 1024])
def test_scatter_partition_equals_the_search(power_law, chunk):
    """Every boundary is written exactly once and equals the merge-path search
    at diagonal min(w·chunk, M + nnz)."""
    for name, rp in _matrices(power_law):
        M, nnz = len(rp) - 1, int(rp[-1])
        part, nwaves = scatter_partition(rp, chunk)
        want = [merge_search(rp, min(w * chunk, M + nnz), M, nnz) for w in range(nwaves + 1)]
        assert part.min() >= 0, name
        assert np.array_equal(part, want), name


def spmv_fixup_model(crow, cval):
    """k_spmv_fixup: a wave per 64 carries, segmented suffix sums by row tag,
    head lanes add their run; the wave holding the head of a run that goes on
    past its lane 63 sums the rest of it."""
    nw = len(crow)
    y = {}
    for base in range(0, nw, 64):
        w = np.arange(base, base + 64)
        r = np.where(w < nw, crow[np.minimum(w, nw - 1)], -2)
        rprev = np.array([-2 if (x == 0 or x > nw) else crow[x - 1] for x in w])
        s = np.where((w < nw) & (r >= 0), cval[np.minimum(w, nw - 1)], 0.0)
        o = 1
        while o < 64:
            t = np.concatenate([s[o:], np.zeros(o)])
            rt = np.concatenate([r[o:], np.full(o, -3)])
            s = s + np.where(rt == r, t, 0.0)
            o <<= 1
        r63 = r[63]
        owner = np.any((r == r63) & (rprev != r63))
        ext = 0.0
        if r63 >= 0 and owner and base + 64 < nw and crow[base + 64] == r63:
            b = base + 64
            while b < nw and crow[b] == r63:
                ext += cval[b]
                b += 1
        for lane in range(64):
            if r[lane] >= 0 and rprev[lane] != r[lane]:
                y[r[lane]] = y.get(r[lane], 0.0) + s[lane] + (ext if r[lane] == r63 else 0.0)
    return y


def test_spmv_fixup_model_sums_every_run_once():
    rng = np.random.default_rng(11)
    crow, row = [], 0
    for _ in range(400):  # runs of 1..300 carries, some waves without a carry
        n = int(rng.choice([1, 1, 2, 3, 63, 64, 65, 130, 300]))
        crow += [row] * n
        if rng.random() < 0.3:
            crow.append(-1)
        row += int(rng.integers(1, 4))
    crow = np.array(crow)
    cval = rng.integers(-9, 10, len(crow)).astype(np.float64)
    y = spmv_fixup_model(crow, cval)
    want = {}
    for r, v in zip(crow, cval):
        if r >= 0:
            want[r] = want.get(r, 0.0) + v
    assert y == want
