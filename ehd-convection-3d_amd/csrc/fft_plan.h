// fft_plan.h -- radix schedule of the LDS Stockham FFTs.
#pragma once

namespace ehd {

struct FftPlan {
    int n;
    int nst;       // number of stages
    int rad[24];   // radix of each stage: 4, 2, 3, 5 butterflies, anything else a direct DFT stage
};

// Radix 4 first, then 2, 3, 5, then the remaining prime factors in ascending order.
inline FftPlan make_fft_plan(int n)
{
    FftPlan p{};
    p.n = n;
    int r = n;
    const int pref[4] = {4, 2, 3, 5};
    for (int i = 0; i < 4; ++i)
        while (r % pref[i] == 0 && p.nst < 24) {
            p.rad[p.nst++] = pref[i];
            r /= pref[i];
        }
    for (int q = 7; r > 1 && p.nst < 24; q += 2)
        while (r % q == 0 && p.nst < 24) {
            p.rad[p.nst++] = q;
            r /= q;
        }
    return p;
}

}  // namespace ehd
