// Host first-fit batching: the reference's CPU batching fallback
// (generate_batches_rsmt, src/batching/generate_batches_cpu.hpp:7-73), which
// the command line runs for `--batching cpu` and uses to validate the GPU
// batches (`--validate-batches 1`).
//
// The reference keeps one X*Y bit-vector per batch and tests a net against
// every open batch in turn.  Here each G-cell holds a bit mask of the batches
// that marked it (one 64-bit plane per 64 batches, planes added as batches
// open), so a net's conflicts are the OR of its points' masks and its batch
// is the lowest clear bit outside the mask of full batches: the work per net
// is |points| * planes plus its marks, independent of how many batches it
// would have tried.  The result is the reference's first fit exactly.
#include <cstdint>
#include <string>
#include <vector>

#include "../csrc/gr_status.hpp"
#include "../../include/grotor_amd.h"

using namespace grotor;

namespace {

struct CellMasks {
    size_t cells;
    std::vector<std::vector<uint64_t>> planes;  // planes[w][cell]: batches 64w .. 64w+63
    std::vector<uint64_t> full;                 // one bit per batch that reached the limit
    explicit CellMasks(size_t n) : cells(n) {}
    void ensure(int batch) {
        const size_t w = size_t(batch) >> 6;
        while (planes.size() <= w) {
            planes.emplace_back(cells, 0);
            full.push_back(0);
        }
    }
    void mark(size_t cell, int batch) { planes[size_t(batch) >> 6][cell] |= uint64_t(1) << (batch & 63); }
};

void check_offsets(int n, const int* off, const char* what) {
    if (off[0] != 0) throw GrError(GR_E_ARG, std::string(what) + " offsets must start at 0");
    for (int k = 0; k < n; ++k)
        if (off[k + 1] < off[k]) throw GrError(GR_E_ARG, std::string(what) + " offsets must not decrease");
}

}  // namespace

extern "C" int gr_batch_plan_host(int X, int Y, int n, const int* pt_off, const int* pts, const int* hs_off,
                                  const int* hsegs, const int* vs_off, const int* vsegs, int limit, int sparsity,
                                  int* color, int* num_batches) {
    return guarded([&] {
        if (X < 1 || Y < 1 || n < 0 || !pt_off || !hs_off || !vs_off || !color || !num_batches)
            throw GrError(GR_E_ARG, "invalid batching arguments");
        if (sparsity < 1) throw GrError(GR_E_ARG, "sparsity must be positive");
        if (limit < 1) throw GrError(GR_E_ARG, "batch limit must be positive");
        const long long XY = (long long)X * Y;
        if (XY >= (1LL << 31)) throw GrError(GR_E_ARG, "grid too large for 2D cell indices");
        check_offsets(n, pt_off, "point");
        check_offsets(n, hs_off, "horizontal segment");
        check_offsets(n, vs_off, "vertical segment");
        if ((pt_off[n] && !pts) || (hs_off[n] && !hsegs) || (vs_off[n] && !vsegs))
            throw GrError(GR_E_ARG, "null geometry array");
        for (int q = 0; q < pt_off[n]; ++q)
            if (pts[q] < 0 || pts[q] >= XY) throw GrError(GR_E_ARG, "batching point outside the grid");
        for (int s = 0; s < hs_off[n]; ++s)
            if (hsegs[2 * s] < 0 || hsegs[2 * s + 1] >= XY) throw GrError(GR_E_ARG, "segment outside the grid");
        for (int s = 0; s < vs_off[n]; ++s)
            if (vsegs[2 * s] < 0 || vsegs[2 * s + 1] >= XY) throw GrError(GR_E_ARG, "segment outside the grid");

        CellMasks m{static_cast<size_t>(XY)};
        std::vector<int> size;
        std::vector<uint64_t> busy;
        for (int k = 0; k < n; ++k) {
            // batches this net cannot join: those marking one of its points, and full ones
            busy.assign(m.full.begin(), m.full.end());
            for (size_t w = 0; w < m.planes.size(); ++w) {
                if (m.full[w] == ~uint64_t(0)) continue;  // 64 full batches: nothing to test
                const uint64_t* plane = m.planes[w].data();
                for (int q = pt_off[k]; q < pt_off[k + 1]; ++q) busy[w] |= plane[pts[q]];
            }
            int b = int(size.size());  // a new batch unless an open one is free
            for (size_t w = 0; w < busy.size(); ++w) {
                const uint64_t freebits = ~busy[w];
                if (!freebits) continue;
                const int cand = int(w * 64) + __builtin_ctzll(freebits);
                if (cand < int(size.size())) b = cand;
                break;
            }
            if (b == int(size.size())) {
                size.push_back(0);
                m.ensure(b);
            }
            color[k] = b;
            if (++size[b] == limit) m.full[size_t(b) >> 6] |= uint64_t(1) << (b & 63);
            // marks: segment cells at the sparsity stride, then each point's
            // stencil {(x-1,y), (x,y), (x+1,y), (x,y-1)}
            const long long hstep = (long long)sparsity * Y;
            for (int s = hs_off[k]; s < hs_off[k + 1]; ++s)
                for (long long p = hsegs[2 * s]; p <= hsegs[2 * s + 1]; p += hstep) m.mark(size_t(p), b);
            for (int s = vs_off[k]; s < vs_off[k + 1]; ++s)
                for (long long p = vsegs[2 * s]; p <= vsegs[2 * s + 1]; p += sparsity) m.mark(size_t(p), b);
            for (int q = pt_off[k]; q < pt_off[k + 1]; ++q) {
                const int x = pts[q] / Y, y = pts[q] % Y;
                m.mark(size_t(pts[q]), b);
                if (x > 0) m.mark(size_t(pts[q] - Y), b);
                if (x + 1 < X) m.mark(size_t(pts[q] + Y), b);
                if (y > 0) m.mark(size_t(pts[q] - 1), b);
            }
        }
        *num_batches = int(size.size());
    });
}
