// Matrix Market reader (host): coordinate files with real, integer or pattern
// values, general, symmetric or skew-symmetric, into canonical CSR.
//
// Reference: the SuiteSparse drivers (suitesparse_spmm.cpp:90-121,
// suitesparse_sdmm.cpp:63-90, suitesparse_spmv.cpp:73-100) skip the `%` lines, read
// "rows cols nnz", scatter every 1-based (i, j, v) entry into a dense matrix
// (a later entry for the same position overwrites an earlier one) and build
// CSR with toCSR, which keeps only nonzero values.  readMtx
// (gcn/util/util.hpp:286-333) reads pattern files as 1.0 and mirrors
// symmetric files.  Here the entries go to CSR without a dense matrix, with
// the same result for every file: last write wins, zeros are dropped, pattern
// entries are 1.0, and symmetric (skew-symmetric: negated) files are
// mirrored, unless ARMSP_MTX_STORED_TRIANGLE asks for the stored entries
// only, as the dense drivers read them.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "armsp_gen.h"

namespace {

struct Entry {
    int r, c;
    int64_t seq;  // write order: a later write to the same position wins
    float v;
};

bool starts_with_ci(const char* s, const char* prefix) {
    for (; *prefix; ++s, ++prefix)
        if (std::tolower((unsigned char)*s) != std::tolower((unsigned char)*prefix)) return false;
    return true;
}

// Parses the file into its entries (mirrored as asked).  0 or an error code.
int parse(const char* path, int flags, int& M, int& K, std::vector<Entry>& out) {
    FILE* f = std::fopen(path, "r");
    if (!f) return 1;
    std::string line;
    char buf[4096];
    bool pattern = false, integer = false, symmetric = false, skew = false;
    bool first = true, have_size = false;
    int64_t declared = 0, seen = 0;
    int rc = 0;
    while (std::fgets(buf, sizeof buf, f)) {
        line.assign(buf);
        while (!line.empty() && line.back() != '\n' && std::fgets(buf, sizeof buf, f)) line += buf;  // long lines
        if (first) {
            first = false;
            if (starts_with_ci(line.c_str(), "%%MatrixMarket")) {
                std::string l = line;
                for (auto& ch : l) ch = (char)std::tolower((unsigned char)ch);
                if (l.find(" matrix ") == std::string::npos || l.find(" coordinate ") == std::string::npos ||
                    l.find(" complex") != std::string::npos || l.find(" hermitian") != std::string::npos) {
                    rc = 2;
                    break;
                }
                pattern = l.find(" pattern") != std::string::npos;
                integer = l.find(" integer") != std::string::npos;
                skew = l.find(" skew-symmetric") != std::string::npos;
                symmetric = skew || l.find(" symmetric") != std::string::npos;
                continue;
            }
        }
        const char* p = line.c_str();
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == '%' || *p == '\n' || *p == '\r' || *p == 0) continue;
        char* end = nullptr;
        if (!have_size) {
            const long long m = std::strtoll(p, &end, 10);
            const long long k = std::strtoll(end, &end, 10);
            const long long n = std::strtoll(end, &end, 10);
            if (end == p || m < 0 || k < 0 || n < 0 || m > 0x7fffffff || k > 0x7fffffff) {
                rc = 3;
                break;
            }
            M = (int)m;
            K = (int)k;
            declared = n;
            have_size = true;
            out.reserve((size_t)n * ((symmetric && !(flags & ARMSP_MTX_STORED_TRIANGLE)) ? 2 : 1));
            continue;
        }
        const long long i = std::strtoll(p, &end, 10);
        const char* q = end;
        const long long j = std::strtoll(q, &end, 10);
        if (end == q) {
            rc = 3;
            break;
        }
        double v = 1.0;
        if (!pattern) {
            q = end;
            v = integer ? (double)std::strtoll(q, &end, 10) : std::strtod(q, &end);
            if (end == q) {
                rc = 3;
                break;
            }
        }
        if (i < 1 || i > M || j < 1 || j > K) {
            rc = 4;
            break;
        }
        const int r = (int)(i - 1), c = (int)(j - 1);
        out.push_back({r, c, 2 * seen, (float)v});
        if (symmetric && !(flags & ARMSP_MTX_STORED_TRIANGLE) && r != c) {
            if (c >= M || r >= K) {
                rc = 4;
                break;
            }
            out.push_back({c, r, 2 * seen + 1, skew ? -(float)v : (float)v});
        }
        ++seen;
    }
    std::fclose(f);
    if (rc) return rc;
    if (!have_size || seen != declared) return 3;
    return 0;
}

}  // namespace

extern "C" int armsp_read_mtx(const char* path, int flags, int* M, int* K, int64_t* nnz, int* rowptr, int* colind,
                              float* val) {
    if (!path || !M || !K || !nnz) return 5;
    if ((rowptr == nullptr) != (colind == nullptr)) return 5;
    int m = 0, k = 0;
    std::vector<Entry> e;
    const int rc = parse(path, flags, m, k, e);
    if (rc) return rc;
    // position order, the last write of each position first among its duplicates
    std::sort(e.begin(), e.end(), [](const Entry& a, const Entry& b) {
        if (a.r != b.r) return a.r < b.r;
        if (a.c != b.c) return a.c < b.c;
        return a.seq > b.seq;
    });
    int64_t n = 0;
    std::vector<int> counts(rowptr ? (size_t)m : 0, 0);
    for (size_t t = 0; t < e.size(); ++t) {
        if (t > 0 && e[t].r == e[t - 1].r && e[t].c == e[t - 1].c) continue;  // overwritten
        if (e[t].v == 0.0f) continue;                                         // toCSR keeps `!= 0`
        if (rowptr) {
            if (n >= *nnz) return 5;  // the arrays were sized by a first call on another file
            colind[n] = e[t].c;
            if (val) val[n] = e[t].v;
            ++counts[e[t].r];
        }
        ++n;
    }
    if (rowptr) {
        if (n != *nnz) return 5;
        rowptr[0] = 0;
        for (int r = 0; r < m; ++r) rowptr[r + 1] = rowptr[r] + counts[r];
    }
    *M = m;
    *K = k;
    *nnz = n;
    return 0;
}
