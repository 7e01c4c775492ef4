/*
 * generators.c -- seeded phase-space generators for the benchmark and test
 * workloads (the role of ls1-MarDyn's src/utils/generator and
 * src/io/*Generator in the reference). Host-only C, no GPU.
 *
 * ls1gen_rsa: random sequential addition. Points are drawn uniformly inside
 * each region in turn and accepted when their periodic minimum-image distance
 * to every accepted point is at least min_sep. A uniform grid of cells with
 * edge >= min_sep holds the accepted points as linked lists, so each trial
 * inspects the 27 surrounding cells only.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

static uint64_t splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double uniform01(uint64_t *s) { return (splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

/* Returns the number of points placed (== sum(counts) on success, fewer if
 * max_trials_per_point * count trials were exhausted for some region). */
long ls1gen_rsa(uint64_t seed, int n_regions, const double *region_lo, const double *region_hi,
                const long *counts, const double *box, double min_sep,
                long max_trials_per_point, double *out) {
    int nc[3];
    double cl[3];
    for (int d = 0; d < 3; ++d) {
        nc[d] = (int)floor(box[d] / min_sep);
        if (nc[d] < 1) nc[d] = 1;
        cl[d] = box[d] / nc[d];
    }
    long ncell = (long)nc[0] * nc[1] * nc[2];
    long total = 0;
    for (int g = 0; g < n_regions; ++g) total += counts[g];
    long *head = (long *)malloc(sizeof(long) * (size_t)ncell);
    long *next = (long *)malloc(sizeof(long) * (size_t)(total + 1));
    if (!head || !next) {
        free(head);
        free(next);
        return -1;
    }
    for (long c = 0; c < ncell; ++c) head[c] = -1;
    double ms2 = min_sep * min_sep;
    uint64_t st = seed * 0x2545F4914F6CDD1Dull + 1;
    long placed = 0;
    for (int g = 0; g < n_regions; ++g) {
        const double *lo = region_lo + 3 * g, *hi = region_hi + 3 * g;
        long want = counts[g], got = 0;
        long trials = 0, max_trials = max_trials_per_point * (want + 1);
        while (got < want && trials < max_trials) {
            ++trials;
            double p[3];
            int ci[3];
            for (int d = 0; d < 3; ++d) {
                p[d] = lo[d] + (hi[d] - lo[d]) * uniform01(&st);
                if (p[d] >= box[d]) p[d] -= box[d];
                if (p[d] < 0) p[d] += box[d];
                ci[d] = (int)(p[d] / cl[d]);
                if (ci[d] >= nc[d]) ci[d] = nc[d] - 1;
            }
            int ok = 1;
            for (int dz = -1; dz <= 1 && ok; ++dz)
                for (int dy = -1; dy <= 1 && ok; ++dy)
                    for (int dx = -1; dx <= 1 && ok; ++dx) {
                        int cx = (ci[0] + dx + nc[0]) % nc[0];
                        int cy = (ci[1] + dy + nc[1]) % nc[1];
                        int cz = (ci[2] + dz + nc[2]) % nc[2];
                        for (long k = head[((long)cz * nc[1] + cy) * nc[0] + cx]; k >= 0; k = next[k]) {
                            double r2 = 0;
                            for (int d = 0; d < 3; ++d) {
                                double dd = p[d] - out[3 * k + d];
                                if (dd > 0.5 * box[d]) dd -= box[d];
                                else if (dd < -0.5 * box[d]) dd += box[d];
                                r2 += dd * dd;
                            }
                            if (r2 < ms2) {
                                ok = 0;
                                break;
                            }
                        }
                    }
            if (!ok) continue;
            for (int d = 0; d < 3; ++d) out[3 * placed + d] = p[d];
            long c = ((long)ci[2] * nc[1] + ci[1]) * nc[0] + ci[0];
            next[placed] = head[c];
            head[c] = placed;
            ++placed;
            ++got;
        }
    }
    free(head);
    free(next);
    return placed;
}
