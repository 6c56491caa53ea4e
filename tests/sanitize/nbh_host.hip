// Host build of the neighbourhood search of rvgp_amd/csrc/nbh_heap.h (the
// code the geo_nbh kernel runs on every lane), for AddressSanitizer and
// UndefinedBehaviorSanitizer runs on the CPU: the reference runs its heap
// with Cython bounds checks off (RVGP/lib/ptu_dijkstra.pyx:217-220).
//
// Input (binary, little endian): int32 n, int32 K, int32 pool, int32 nnz,
// rowptr int32[n+1], cols int32[nnz], w float64[nnz].
// Output: int32 status (0 ok, 1 pool overflow, 2 short component) then
// int32 nbh[n*(K+1)].  Each lane's buffers are separate heap allocations of
// exactly the size the kernel's plan gives them, so an access past any of
// them is reported.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../rvgp_amd/csrc/nbh_heap.h"

using namespace rvgp;

static bool read_all(FILE* f, void* p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[4];
    if (!read_all(f, hdr, sizeof hdr)) return 2;
    const int n = hdr[0], K = hdr[1], pool = hdr[2], nnz = hdr[3];
    std::vector<int> rowptr(n + 1), cols(nnz);
    std::vector<double> w(nnz);
    if (!read_all(f, rowptr.data(), 4 * (size_t)(n + 1)) || !read_all(f, cols.data(), 4 * (size_t)nnz) ||
        !read_all(f, w.data(), 8 * (size_t)nnz))
        return 2;
    fclose(f);
    int hsize = 1;
    while (hsize < 2 * pool) hsize <<= 1;
    NbhLane L;
    L.N = static_cast<HNode*>(malloc(sizeof(HNode) * (size_t)pool));
    L.tab = static_cast<int2*>(malloc(sizeof(int2) * (size_t)hsize));
    L.hposv = static_cast<int*>(malloc(sizeof(int) * (size_t)pool));
    L.roots = static_cast<int*>(malloc(sizeof(int) * (size_t)kMaxRank));
    L.pool = pool;
    L.mask = (unsigned)hsize - 1u;
    for (int q = 0; q < hsize; ++q) L.tab[q] = make_int2(-1, -1);
    std::vector<int> nbh((size_t)n * (K + 1), -1);
    int status = 0;
    std::vector<int> out(K + 1);
    for (int src = 0; src < n; ++src) {
        int scanned = 0;
        if (!nbh_search(rowptr.data(), cols.data(), w.data(), K, src, L, out.data(), &scanned)) {
            status = 1;
            break;
        }
        if (scanned < K + 1) status = status > 2 ? status : 2;
        for (int q = 0; q < scanned; ++q) nbh[(size_t)src * (K + 1) + q] = out[q];
    }
    free(L.N);
    free(L.tab);
    free(L.hposv);
    free(L.roots);
    FILE* g = fopen(argv[2], "wb");
    if (!g) return 2;
    fwrite(&status, 4, 1, g);
    fwrite(nbh.data(), 4, nbh.size(), g);
    fclose(g);
    return 0;
}
