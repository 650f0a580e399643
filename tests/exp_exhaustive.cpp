// Exhaustive host check of grotor_amd/csrc/gr_exp.hpp: for every binary32
// argument y, the fast form gr_expf_fast(y) returns the bits of the
// reference form gr_expf(y); for every y in [-88, 89] and every step delta
// whose base x0 = y - delta passes gr_expf_step's conditions, the derived
// value gr_expf_step(y, x0, v0) returns them as well.  Also counts the raw
// table values that would round differently without the tie fallback (the
// fallback is needed exactly there) and samples the relative error of both
// binary64 forms against expl.
//
// Built by tests/test_exp_exhaustive.py with hipcc for the host only, with
// the library's -ffp-contract=off.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "gr_exp.hpp"

using namespace grotor;

namespace {

struct Tally {
    unsigned long long mismatch = 0, step_mismatch = 0, raw_mismatch = 0, ties = 0, steps = 0;
    long double err_table = 0, err_horner = 0;
};

unsigned bits_of(float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    return u;
}

float float_of(unsigned u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

void scan(unsigned long long begin, unsigned long long end, Tally* t) {
    const double deltas[4] = {0.25, 0.5, 0.75, 1.5};
    for (unsigned long long u = begin; u < end; ++u) {
        const float y = float_of((unsigned)u);
        const float ref = gr_expf(y);
        double v;
        const float fast = gr_expf_fast(y, &v);
        if (bits_of(ref) != bits_of(fast)) {
            if (t->mismatch < 8) std::printf("mismatch y=%a ref=%a fast=%a\n", y, ref, fast);
            ++t->mismatch;
        }
        if (!(y >= -88.0f && y <= 89.0f)) continue;
        const double table = gr_exp_table((double)y);
        if (gr_exp_near_tie(table)) ++t->ties;
        if (bits_of(gr_exp_round(table)) != bits_of(ref)) ++t->raw_mismatch;
        if ((u & 63) == 0) {
            const long double e = expl((long double)y);
            const long double et = fabsl(((long double)table - e) / e);
            const long double eh = fabsl(((long double)gr_exp_horner((double)y) - e) / e);
            if (et > t->err_table) t->err_table = et;
            if (eh > t->err_horner) t->err_horner = eh;
        }
        for (double d : deltas) {
            const double b = (double)y - d;
            const float x0 = (float)b;
            if ((double)x0 != b) continue;  // y - delta must round to x0 itself
            ++t->steps;
            double v0, vs;
            gr_expf_fast(x0, &v0);
            const float s = gr_expf_step(y, x0, v0, &vs);
            if (bits_of(s) != bits_of(ref)) {
                if (t->step_mismatch < 8) std::printf("step mismatch y=%a x0=%a ref=%a got=%a\n", y, x0, ref, s);
                ++t->step_mismatch;
            }
        }
    }
}

}  // namespace

int main(int argc, char** argv) {
    int threads = argc > 1 ? std::atoi(argv[1]) : 8;
    if (threads < 1) threads = 1;
    const unsigned long long total = 1ull << 32;
    std::vector<Tally> tally(threads);
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; ++i)
        pool.emplace_back(scan, total * i / threads, total * (i + 1) / threads, &tally[i]);
    for (auto& th : pool) th.join();
    Tally s;
    for (const Tally& t : tally) {
        s.mismatch += t.mismatch;
        s.step_mismatch += t.step_mismatch;
        s.raw_mismatch += t.raw_mismatch;
        s.ties += t.ties;
        s.steps += t.steps;
        if (t.err_table > s.err_table) s.err_table = t.err_table;
        if (t.err_horner > s.err_horner) s.err_horner = t.err_horner;
    }
    std::printf("mismatch %llu step_mismatch %llu raw_mismatch %llu ties %llu steps %llu "
                "err_table %.3Le err_horner %.3Le\n",
                s.mismatch, s.step_mismatch, s.raw_mismatch, s.ties, s.steps, s.err_table,
                s.err_horner);
    return 0;
}
