/*
 * CPU matrix-free Rydberg H·v — TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Restates the reference's channel semantics (tomisnot/energy-level
 * src/core/manybody/channels.py:58-77 diagonal, :91-132 flip channels,
 * :210-221 matvec) for the Rydberg model: Ω flips, +delta·n_i, V·n_i n_j.
 * Site i is bit (n-1-i) of a code (reference basis.py:6-9).  Only tests/,
 * the smoke check and bench.py's cpu_baseline leg load this library.
 *
 * Full basis: y[t] = d[t] x[t] + Σ_i a_i(b_i(t)) x[t ^ (1<<(n-1-i))], where
 * a_i = Ω_i/2 when b_i(t)=1 (g→r) and conj(Ω_i)/2 when b_i(t)=0.
 * Chain-blockade (Fibonacci) sector: the basis is all codes with no two
 * adjacent sites excited, in ascending order; rank(code) = Σ_{b_i=1} F(n-i+1).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#ifdef _OPENMP
#include <omp.h>
#endif

int ref_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void ref_set_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

/* diag[k] for codes[k] (codes == NULL → full basis k = code). */
void ref_diag(int n, const double* delta, int n_pairs, const int32_t* pi,
              const int32_t* pj, const double* V, const int64_t* codes,
              int64_t dim, double* diag) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < dim; ++k) {
        int64_t s = codes ? codes[k] : k;
        double d = 0.0;
        for (int i = 0; i < n; ++i)
            if ((s >> (n - 1 - i)) & 1) d += delta[i];
        for (int p = 0; p < n_pairs; ++p)
            if (((s >> (n - 1 - pi[p])) & 1) && ((s >> (n - 1 - pj[p])) & 1)) d += V[p];
        diag[k] = d;
    }
}

void ref_hv_full_f64(int n, const double* om_re, const double* diag,
                     const double* x, double* y) {
    const int64_t D = (int64_t)1 << n;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < D; ++t) {
        double acc = diag[t] * x[t];
        for (int i = 0; i < n; ++i) {
            if (om_re[i] != 0.0) acc += 0.5 * om_re[i] * x[t ^ ((int64_t)1 << (n - 1 - i))];
        }
        y[t] = acc;
    }
}

void ref_hv_full_f32(int n, const double* om_re, const double* diag,
                     const float* x, float* y) {
    const int64_t D = (int64_t)1 << n;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < D; ++t) {
        float acc = (float)diag[t] * x[t];
        for (int i = 0; i < n; ++i) {
            if (om_re[i] != 0.0) acc += (float)(0.5 * om_re[i]) * x[t ^ ((int64_t)1 << (n - 1 - i))];
        }
        y[t] = acc;
    }
}

/* complex128 interleaved (re, im) */
void ref_hv_full_c128(int n, const double* om_re, const double* om_im,
                      const double* diag, const double* x, double* y) {
    const int64_t D = (int64_t)1 << n;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < D; ++t) {
        double ar = diag[t] * x[2 * t], ai = diag[t] * x[2 * t + 1];
        for (int i = 0; i < n; ++i) {
            if (om_re[i] == 0.0 && om_im[i] == 0.0) continue;
            int64_t s = t ^ ((int64_t)1 << (n - 1 - i));
            int bt = (int)((t >> (n - 1 - i)) & 1);
            double cr = 0.5 * om_re[i], ci = bt ? 0.5 * om_im[i] : -0.5 * om_im[i];
            double xr = x[2 * s], xi = x[2 * s + 1];
            ar += cr * xr - ci * xi;
            ai += cr * xi + ci * xr;
        }
        y[2 * t] = ar;
        y[2 * t + 1] = ai;
    }
}

static int64_t fib(int k) {
    int64_t a = 0, b = 1;
    for (int i = 0; i < k; ++i) { int64_t c = a + b; a = b; b = c; }
    return a;
}

/* code of the rank-th chain-blockade state (greedy Zeckendorf). */
int64_t ref_fib_unrank(int n, int64_t r) {
    int64_t code = 0;
    int prev = 0;
    for (int i = 0; i < n; ++i) {
        int64_t f = fib(n - i + 1);
        if (!prev && r >= f) {
            code |= (int64_t)1 << (n - 1 - i);
            r -= f;
            prev = 1;
        } else {
            prev = 0;
        }
    }
    return code;
}

int64_t ref_fib_dim(int n) { return fib(n + 2); }

/* all codes of the chain sector in ascending order */
void ref_fib_codes(int n, int64_t* codes) {
    const int64_t D = fib(n + 2);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < D; ++r) codes[r] = ref_fib_unrank(n, r);
}

/* chain-sector H·v (FP64 real); diag over the sector codes */
void ref_hv_fib_f64(int n, const double* om_re, const double* diag,
                    const int64_t* codes, const double* x, double* y) {
    const int64_t D = fib(n + 2);
    int64_t F[64];
    for (int i = 0; i < n; ++i) F[i] = fib(n - i + 1);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < D; ++r) {
        int64_t c = codes[r];
        double acc = diag[r] * x[r];
        for (int i = 0; i < n; ++i) {
            if (om_re[i] == 0.0) continue;
            int b = (int)((c >> (n - 1 - i)) & 1);
            if (b) {
                acc += 0.5 * om_re[i] * x[r - F[i]];
            } else {
                int left = i > 0 ? (int)((c >> (n - i)) & 1) : 0;
                int right = i < n - 1 ? (int)((c >> (n - 2 - i)) & 1) : 0;
                if (!left && !right) acc += 0.5 * om_re[i] * x[r + F[i]];
            }
        }
        y[r] = acc;
    }
}

void ref_hv_fib_f32(int n, const double* om_re, const double* diag,
                    const int64_t* codes, const float* x, float* y) {
    const int64_t D = fib(n + 2);
    int64_t F[64];
    for (int i = 0; i < n; ++i) F[i] = fib(n - i + 1);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < D; ++r) {
        int64_t c = codes[r];
        float acc = (float)diag[r] * x[r];
        for (int i = 0; i < n; ++i) {
            if (om_re[i] == 0.0) continue;
            float a = (float)(0.5 * om_re[i]);
            int b = (int)((c >> (n - 1 - i)) & 1);
            if (b) {
                acc += a * x[r - F[i]];
            } else {
                int left = i > 0 ? (int)((c >> (n - i)) & 1) : 0;
                int right = i < n - 1 ? (int)((c >> (n - 2 - i)) & 1) : 0;
                if (!left && !right) acc += a * x[r + F[i]];
            }
        }
        y[r] = acc;
    }
}

/* rows [t0, t1) of the full-basis FP64 H·v; diag_rows[k] is the diagonal of
 * code t0 + k (used to time a bounded sample of a large workload) */
void ref_hv_full_f64_rows(int n, const double* om_re, const double* diag_rows,
                          const double* x, double* y, int64_t t0, int64_t t1) {
#pragma omp parallel for schedule(static)
    for (int64_t t = t0; t < t1; ++t) {
        double acc = diag_rows[t - t0] * x[t];
        for (int i = 0; i < n; ++i) {
            if (om_re[i] != 0.0) acc += 0.5 * om_re[i] * x[t ^ ((int64_t)1 << (n - 1 - i))];
        }
        y[t - t0] = acc;
    }
}

/* ---- blockade graphs (reference basis.py:24-47) ----------------------------
 * Independent sets of a graph given by per-site neighbour masks (nb[i] has
 * bit j set when sites i and j are joined), in the reference's depth-first
 * order: site 0 first, "empty" before "excited", so codes come out
 * ascending.  The recursion is unrolled into an explicit stack.
 * Returns the number of sets; writes at most cap codes (codes may be NULL). */
int64_t ref_indep_codes(int n, const uint64_t* nb, int64_t* codes, int64_t cap) {
    int64_t count = 0;
    if (n <= 0) {
        if (codes && cap > 0) codes[0] = 0;
        return 1;
    }
    /* choice[s] = 0: site s empty taken, 1: excited taken */
    int choice[64];
    uint64_t occ[65];   /* occupied sites (by site index) before site s */
    int64_t state[65];
    int s = 0;
    occ[0] = 0;
    state[0] = 0;
    choice[0] = -1;
    while (s >= 0) {
        if (s == n) {
            if (codes && count < cap) codes[count] = state[n];
            ++count;
            --s;
            continue;
        }
        if (choice[s] == -1) {             /* empty first */
            choice[s] = 0;
            occ[s + 1] = occ[s];
            state[s + 1] = state[s];
            if (s + 1 < n) choice[s + 1] = -1;
            ++s;
            continue;
        }
        if (choice[s] == 0 && !(nb[s] & occ[s])) {   /* then excited, if allowed */
            choice[s] = 1;
            occ[s + 1] = occ[s] | ((uint64_t)1 << s);
            state[s + 1] = state[s] | ((int64_t)1 << (n - 1 - s));
            if (s + 1 < n) choice[s + 1] = -1;
            ++s;
            continue;
        }
        --s;                                /* both branches done */
    }
    return count;
}

/* y = H x on a sorted code list (FP64 real Ω): every flip of site i whose
 * target code is in the list (binary search) contributes Ω_i/2 x[target];
 * sources outside the list are dropped (the projected operator, as the
 * reference's restricted-basis ChannelKernel, channels.py:177-221). */
void ref_hv_codes_f64(int n, const double* om_re, const double* diag,
                      const int64_t* codes, int64_t dim, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < dim; ++r) {
        const int64_t c = codes[r];
        double acc = diag[r] * x[r];
        for (int i = 0; i < n; ++i) {
            if (om_re[i] == 0.0) continue;
            const int64_t s = c ^ ((int64_t)1 << (n - 1 - i));
            int64_t lo = 0, hi = dim - 1;
            while (lo <= hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (codes[mid] == s) { acc += 0.5 * om_re[i] * x[mid]; break; }
                if (codes[mid] < s) lo = mid + 1; else hi = mid - 1;
            }
        }
        y[r] = acc;
    }
}
