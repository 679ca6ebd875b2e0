// Host check of the autotuner's admission rule (ssx::admit_candidates) on
// every assignment of ok flags and digests (3 digest values) to 1..6
// candidates, by the properties the rule must have rather than a second copy
// of it: admitted candidates are ok and share one digest; every ok candidate
// with that digest is admitted; no group of agreeing ok candidates is larger;
// a lone ok candidate is admitted; with two or more ok candidates something
// is admitted iff two of them agree.  Prints the number of failed cases; exit
// status 1 if any.
#include <cstdio>

#include "ss_admit.hpp"

int main() {
    long cases = 0, failed = 0;
    for (int n = 1; n <= 6; ++n) {
        int pow3 = 1;
        for (int i = 0; i < n; ++i) pow3 *= 3;
        for (int okm = 0; okm < (1 << n); ++okm) {
            for (int dm = 0; dm < pow3; ++dm) {
                bool ok[6], adm[6];
                uint64_t dig[6];
                int t = dm;
                for (int i = 0; i < n; ++i) {
                    ok[i] = (okm >> i) & 1;
                    dig[i] = 1000u + (uint64_t)(t % 3);
                    t /= 3;
                }
                const int got = ssx::admit_candidates(ok, dig, n, adm);
                int n_ok = 0, n_adm = 0, group_max = 0, adm_digest = -1;
                bool bad = false;
                for (int i = 0; i < n; ++i) {
                    n_ok += ok[i];
                    if (!ok[i]) continue;
                    int c = 0;
                    for (int j = 0; j < n; ++j) c += ok[j] && dig[j] == dig[i];
                    if (c > group_max) group_max = c;
                }
                for (int i = 0; i < n; ++i) {
                    if (!adm[i]) continue;
                    ++n_adm;
                    if (!ok[i]) bad = true;
                    if (adm_digest < 0) adm_digest = (int)dig[i];
                    else if ((int)dig[i] != adm_digest) bad = true;
                }
                if (got != n_adm) bad = true;
                if (n_adm) {
                    for (int i = 0; i < n; ++i)
                        if (ok[i] && (int)dig[i] == adm_digest && !adm[i]) bad = true;
                    if (n_adm < group_max) bad = true;
                }
                if (n_ok == 1 && n_adm != 1) bad = true;
                if (n_ok >= 2 && (n_adm > 0) != (group_max >= 2)) bad = true;
                if (n_ok == 0 && n_adm != 0) bad = true;
                ++cases;
                failed += bad;
            }
        }
    }
    std::printf("%ld cases, %ld failed\n", cases, failed);
    return failed ? 1 : 0;
}
