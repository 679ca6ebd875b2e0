// ss_admit.hpp -- admission rule of the autotuner (host only, no HIP).
//
// The reference admits a (T, threads) build to timing only after it passes
// the oracle suites with zero mismatches on every route
// (harness/configuration_study/run_config_study_v0_8.ps1:57-68; the guards of
// sparse_config_benchmark_v0_8.cu:196-231).  The library has no fixtures at
// run time, so a candidate is admitted when (a) its runs leave the device
// invariant word at zero and (b) its masks on the caller's images agree bit for
// bit with those of the largest group of candidates that agree with each other
// (the candidates decompose the images differently -- other tile sides, other
// seams -- so agreement of independent decompositions is the check).
#pragma once
#include <stdint.h>

namespace ssx {

// ok[i]: candidate i ran with a zero invariant word; digest[i]: digest of its
// masks.  Sets admitted[i] and returns how many candidates are admitted.  The
// admitted group is the most common digest among the ok candidates (ties: the
// one seen first).  With two or more ok candidates the group must hold at
// least two of them, or nothing is admitted: one candidate that agrees with
// no other is no evidence of correctness.
inline int admit_candidates(const bool* ok, const uint64_t* digest, int n, bool* admitted) {
    int n_ok = 0, best = -1, best_count = 0;
    for (int i = 0; i < n; ++i) {
        admitted[i] = false;
        if (!ok[i]) continue;
        ++n_ok;
        int count = 0;
        for (int j = 0; j < n; ++j) count += (ok[j] && digest[j] == digest[i]) ? 1 : 0;
        if (count > best_count) {
            best_count = count;
            best = i;
        }
    }
    if (best < 0 || (n_ok >= 2 && best_count < 2)) return 0;
    int n_adm = 0;
    for (int i = 0; i < n; ++i) {
        admitted[i] = ok[i] && digest[i] == digest[best];
        n_adm += admitted[i] ? 1 : 0;
    }
    return n_adm;
}

}  // namespace ssx
