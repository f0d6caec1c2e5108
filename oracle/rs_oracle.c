/*
 * rs_oracle.c — CPU FP64 restatement of CELERIS's Rayleigh–Sommerfeld far-field sum.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing
 * else.  The product library (celeris_amd/) never links, loads or calls it.
 *
 * What it restates (reference file:line):
 *   - propagate_pillars, the FP64 CPU loop       src/celeris/analysis/focal.cpp:205-238
 *       fx = cx - W + i*step, fy = cy - W + j*step, step = 2W/(n-1)   (:223-227)
 *       R  = sqrt((fx-px)^2 + (fy-py)^2 + z^2)                       (:230-231)
 *       E += t * polar(1/R, k*R), pillars in list order             (:232)
 *       I  = |E|^2                                                   (:234)
 *   - parallel_rows (row striding over workers)  src/celeris/analysis/focal.cpp:19-29
 *   - the y=0 caustic plane (zx)                 src/celeris/cuda/propagate.cu:97-125, 205-206
 *       fx = xmin + ix*xstep, z = zmin + iz*zstep, dy = y_pillar
 *   - the adjoint g_q = sum_p conj(G_pq) u_p has no reference counterpart; it is
 *     the conjugate transpose of the sum above, evaluated with the same G_pq.
 *   - the opt-in Rayleigh-Sommerfeld I kernel (plane kernel 1) has no reference
 *     counterpart either: G = exp(ikR)/R (z/R)(1/R - ik) = -d/dz [exp(ikR)/R].
 *   - the pupil autocorrelation of analyze_mtf     src/celeris/analysis/mtf.cpp:40-49
 *       acc(s) = sum_iy sum_{ix+s<n} P[iy][ix] conj(P[iy][ix+s]), rows then columns
 *   - the fill jitter of analyze_tolerance         src/celeris/analysis/tolerance.cpp:20-31
 *       fill = clamp(fill + N(0, sigma_fill), 0.05, 0.95) per map cell, from a
 *       std::mt19937 seeded seed + trial*7919 + unsigned(sigma*13) and
 *       std::normal_distribution<double>, restated from the published algorithms below
 *
 * Pixel geometry is expressed as an affine "plane": pixel (i, j) sits at
 *   fx = x0 + i*sx,  fy = y0 + j*sy,  z = z0 + j*sz
 * which reproduces the psf window (x0 = cx-W, sx = step, y0 = cy-W, sy = step,
 * sz = 0) and the zx slice (x0 = xmin, sx = xstep, y0 = sy = 0, z0 = zmin,
 * sz = zstep) with the reference's own evaluation order.
 *
 * Pinning: tests/test_oracle.py checks these functions against the reference's
 * own compiled CPU path (oracle/_ref, built from the reference sources), against
 * committed golden vectors produced by that build, against closed forms (single
 * pillar, in-phase focus, linearity, mirror symmetry), against the long-double
 * brute force below, and the adjoint against the dot-product identity and
 * finite differences of the forward.  The RS-I kernel is pinned by the on-axis
 * single-pillar closed form exp(ikz)(1/z - ik)/z, by central differences in z of
 * the reference kernel (G_RSI = -dG/dz), and by the long-double brute force.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

typedef struct orc_plane {
    double k;
    double x0, sx;
    double y0, sy;
    double z0, sz;
    int32_t nx, ny;
    int32_t tset;
    int32_t kernel; /* 0: exp(ikR)/R; 1: RS-I, exp(ikR)/R (z/R)(1/R - ik) */
    int64_t offset;
} orc_plane;

static int online_cpus(void) {
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

/* Pair kernel of the reference: G = polar(1/R, k R), R from the plane point
 * (focal.cpp:230-232).  kernel 1, which the reference does not have: the
 * Rayleigh-Sommerfeld I kernel G = exp(ikR)/R (z/R)(1/R - ik), evaluated as
 * polar(1/R, k R) times the factor (z/R^2)(1/R - ik). */
static inline void pair_g(double fx, double fy, double z, double k, double qx,
                          double qy, int kernel, double *gr, double *gi) {
    const double dx = fx - qx, dy = fy - qy;
    const double R = sqrt(dx * dx + dy * dy + z * z);
    const double rho = 1.0 / R, th = k * R;
    const double pr = rho * cos(th), pi_ = rho * sin(th);
    if (kernel == 1) {
        const double a = z * rho;                 /* z / R */
        const double fr = a * rho, fi = -a * k;   /* (z/R)(1/R - ik) */
        *gr = pr * fr - pi_ * fi;
        *gi = pr * fi + pi_ * fr;
    } else {
        *gr = pr;
        *gi = pi_;
    }
}

/* ------------------------------------------------------------------------- */
/* Row tables shared by forward and adjoint.                                 */

typedef struct {
    int64_t total_rows;
    int64_t *row_plane; /* global row -> plane index */
    int32_t *row_j;     /* global row -> row within plane */
} row_table;

static int make_rows(const orc_plane *planes, int nplanes, row_table *rt) {
    int64_t total = 0;
    for (int p = 0; p < nplanes; ++p) {
        if (planes[p].nx < 0 || planes[p].ny < 0) return -1;
        total += planes[p].ny;
    }
    rt->total_rows = total;
    rt->row_plane = (int64_t *)malloc(sizeof(int64_t) * (size_t)(total + 1));
    rt->row_j = (int32_t *)malloc(sizeof(int32_t) * (size_t)(total + 1));
    if (!rt->row_plane || !rt->row_j) {
        free(rt->row_plane); free(rt->row_j);
        return -2;
    }
    int64_t r = 0;
    for (int p = 0; p < nplanes; ++p)
        for (int j = 0; j < planes[p].ny; ++j) { rt->row_plane[r] = p; rt->row_j[r] = j; ++r; }
    return 0;
}

static void free_rows(row_table *rt) { free(rt->row_plane); free(rt->row_j); }

/* ------------------------------------------------------------------------- */
/* Forward: one complex field value per pixel.                               */

typedef struct {
    const double *px, *py, *t; /* t: interleaved (re, im), nsets x npil */
    int npil;
    const orc_plane *planes;
    double *field;     /* interleaved complex, or NULL */
    double *intensity; /* or NULL */
    const row_table *rt;
    int row_stride;    /* evaluate global rows r with r % row_stride == 0 */
    int workers, w;
    int64_t pairs;     /* pairs evaluated by this worker */
} fwd_job;

static void fwd_row(fwd_job *J, int64_t r) {
    const orc_plane *P = &J->planes[J->rt->row_plane[r]];
    const int j = J->rt->row_j[r];
    const double fy = P->y0 + j * P->sy;
    const double z = P->z0 + j * P->sz;
    const double k = P->k;
    const double *t = J->t + (size_t)2 * (size_t)P->tset * (size_t)J->npil;
    for (int i = 0; i < P->nx; ++i) {
        const double fx = P->x0 + i * P->sx;
        double er = 0.0, ei = 0.0;
        for (int q = 0; q < J->npil; ++q) {
            double pr, pim;
            pair_g(fx, fy, z, k, J->px[q], J->py[q], P->kernel, &pr, &pim);
            const double tr = t[2 * q], ti = t[2 * q + 1];
            er += tr * pr - ti * pim;
            ei += tr * pim + ti * pr;
        }
        const int64_t o = P->offset + (int64_t)j * P->nx + i;
        if (J->field) { J->field[2 * o] = er; J->field[2 * o + 1] = ei; }
        if (J->intensity) J->intensity[o] = er * er + ei * ei;
    }
    J->pairs += (int64_t)P->nx * J->npil;
}

static void *fwd_worker(void *arg) {
    fwd_job *J = (fwd_job *)arg;
    /* parallel_rows: worker w takes rows w, w+W, ... (focal.cpp:26).  A row sample
     * (row_stride > 1) is striped the same way over the sampled rows, so every worker
     * stays busy whatever the stride and the worker count have in common. */
    for (int64_t s = J->w; s * J->row_stride < J->rt->total_rows; s += J->workers)
        fwd_row(J, s * J->row_stride);
    return NULL;
}

/* Forward sum over all planes.  Rows of all planes are striped over `threads`
 * workers (threads <= 0: one per online CPU).  row_stride > 1 evaluates only
 * every row_stride-th global row (a bounded timing sample); other rows are left
 * untouched.  Returns the number of pillar-pixel pairs evaluated, < 0 on error. */
int64_t orc_forward(const double *px, const double *py, const double *t, int npil,
                    const orc_plane *planes, int nplanes, double *field,
                    double *intensity, int threads, int row_stride) {
    if (npil < 0 || nplanes < 0) return -1;
    if (row_stride < 1) row_stride = 1;
    row_table rt;
    if (make_rows(planes, nplanes, &rt)) return -2;
    if (threads <= 0) threads = online_cpus();
    int workers = threads;
    if (workers > rt.total_rows) workers = (int)(rt.total_rows > 0 ? rt.total_rows : 1);
    fwd_job *jobs = (fwd_job *)calloc((size_t)workers, sizeof(fwd_job));
    pthread_t *th = (pthread_t *)calloc((size_t)workers, sizeof(pthread_t));
    for (int w = 0; w < workers; ++w) {
        fwd_job J = {px, py, t, npil, planes, field, intensity, &rt, row_stride,
                     workers, w, 0};
        jobs[w] = J;
    }
    for (int w = 1; w < workers; ++w) pthread_create(&th[w], NULL, fwd_worker, &jobs[w]);
    fwd_worker(&jobs[0]);
    int64_t pairs = 0;
    for (int w = 1; w < workers; ++w) pthread_join(th[w], NULL);
    for (int w = 0; w < workers; ++w) pairs += jobs[w].pairs;
    free(jobs); free(th); free_rows(&rt);
    return pairs;
}

/* ------------------------------------------------------------------------- */
/* Adjoint: g[s][q] = sum over planes with tset == s, over their pixels p, of */
/* conj(G_pq) * u_p.  With u = dL/dRe(E) + i dL/dIm(E) this is                */
/* dL/dRe(t) + i dL/dIm(t) for any real loss L of the forward field E.        */

typedef struct {
    const double *px, *py;
    int npil, nsets;
    const orc_plane *planes;
    int nplanes;
    const double *u; /* interleaved complex per pixel */
    double *g;       /* interleaved complex, nsets x npil */
    int pillar_stride;
    int workers, w;
    int64_t pairs;
} adj_job;

static void *adj_worker(void *arg) {
    adj_job *J = (adj_job *)arg;
    /* sampled pillars q = i * pillar_stride, striped over workers by i */
    for (int64_t i = J->w; i * J->pillar_stride < J->npil; i += J->workers) {
        const int q = (int)(i * J->pillar_stride);
        const double qx = J->px[q], qy = J->py[q];
        for (int s = 0; s < J->nsets; ++s) {
            double gr = 0.0, gi = 0.0;
            for (int p = 0; p < J->nplanes; ++p) {
                const orc_plane *P = &J->planes[p];
                if (P->tset != s) continue;
                for (int j = 0; j < P->ny; ++j) {
                    const double fy = P->y0 + j * P->sy;
                    const double z = P->z0 + j * P->sz;
                    const double *u = J->u + 2 * (P->offset + (int64_t)j * P->nx);
                    for (int i = 0; i < P->nx; ++i) {
                        const double fx = P->x0 + i * P->sx;
                        double cr, ci;
                        pair_g(fx, fy, z, P->k, qx, qy, P->kernel, &cr, &ci);
                        const double ur = u[2 * i], ui = u[2 * i + 1];
                        /* conj(G) * u = (cr - i ci)(ur + i ui) */
                        gr += cr * ur + ci * ui;
                        gi += cr * ui - ci * ur;
                    }
                    J->pairs += P->nx;
                }
            }
            J->g[2 * ((size_t)s * J->npil + q)] = gr;
            J->g[2 * ((size_t)s * J->npil + q) + 1] = gi;
        }
    }
    return NULL;
}

/* Adjoint of orc_forward.  Pillars are striped over workers; pillar_stride > 1
 * evaluates only pillars q with q % pillar_stride == 0 (others untouched).
 * Returns the number of pairs evaluated, < 0 on error. */
int64_t orc_adjoint(const double *px, const double *py, int npil, int nsets,
                    const orc_plane *planes, int nplanes, const double *u, double *g,
                    int threads, int pillar_stride) {
    if (npil < 0 || nsets < 1 || nplanes < 0) return -1;
    if (pillar_stride < 1) pillar_stride = 1;
    if (threads <= 0) threads = online_cpus();
    int workers = threads;
    if (workers > npil) workers = npil > 0 ? npil : 1;
    adj_job *jobs = (adj_job *)calloc((size_t)workers, sizeof(adj_job));
    pthread_t *th = (pthread_t *)calloc((size_t)workers, sizeof(pthread_t));
    for (int w = 0; w < workers; ++w) {
        adj_job J = {px, py, npil, nsets, planes, nplanes, u, g, pillar_stride,
                     workers, w, 0};
        jobs[w] = J;
    }
    for (int w = 1; w < workers; ++w) pthread_create(&th[w], NULL, adj_worker, &jobs[w]);
    adj_worker(&jobs[0]);
    int64_t pairs = 0;
    for (int w = 1; w < workers; ++w) pthread_join(th[w], NULL);
    for (int w = 0; w < workers; ++w) pairs += jobs[w].pairs;
    free(jobs); free(th);
    return pairs;
}

/* ------------------------------------------------------------------------- */
/* Extended-precision brute force (x87 long double, 64-bit mantissa), single  */
/* threaded, for tiny inputs: pins the FP64 restatement's rounding.  Inputs   */
/* are the same doubles; geometry and sum are carried in long double.         */

int orc_forward_ld(const double *px, const double *py, const double *t, int npil,
                   const orc_plane *planes, int nplanes, double *field) {
    for (int p = 0; p < nplanes; ++p) {
        const orc_plane *P = &planes[p];
        const double *tt = t + (size_t)2 * (size_t)P->tset * (size_t)npil;
        for (int j = 0; j < P->ny; ++j) {
            const long double fy = (long double)P->y0 + (long double)j * P->sy;
            const long double z = (long double)P->z0 + (long double)j * P->sz;
            for (int i = 0; i < P->nx; ++i) {
                const long double fx = (long double)P->x0 + (long double)i * P->sx;
                long double er = 0.0L, ei = 0.0L;
                for (int q = 0; q < npil; ++q) {
                    const long double dx = fx - px[q], dy = fy - py[q];
                    const long double R = sqrtl(dx * dx + dy * dy + z * z);
                    const long double th = (long double)P->k * R;
                    long double pr = cosl(th) / R, pim = sinl(th) / R;
                    if (P->kernel == 1) {
                        const long double fr = z / (R * R), fi = -(long double)P->k * z / R;
                        const long double gr = pr * fr - pim * fi;
                        pim = pr * fi + pim * fr;
                        pr = gr;
                    }
                    er += tt[2 * q] * pr - tt[2 * q + 1] * pim;
                    ei += tt[2 * q] * pim + tt[2 * q + 1] * pr;
                }
                const int64_t o = P->offset + (int64_t)j * P->nx + i;
                field[2 * o] = (double)er;
                field[2 * o + 1] = (double)ei;
            }
        }
    }
    return 0;
}

int orc_num_cpus(void) { return online_cpus(); }

/* Pupil autocorrelation along x for shifts 0..nshift-1 (mtf.cpp:40-49), in the
 * reference's loop order: P complex (interleaved) n x n, P0 real n x n.  Writes
 * acc[2s], acc[2s+1] = sum P conj(P shifted by s) and acc0[s] = sum P0 P0(shifted). */
void orc_mtf_autocorr(const double *P, const double *P0, int n, int nshift, double *acc,
                      double *acc0) {
    for (int s = 0; s < nshift; ++s) {
        double ar = 0.0, ai = 0.0, a0 = 0.0;
        for (int iy = 0; iy < n; ++iy) {
            const double *row = P + 2 * (int64_t)iy * n;
            const double *row0 = P0 + (int64_t)iy * n;
            for (int ix = 0; ix + s < n; ++ix) {
                const double px = row[2 * ix], py = row[2 * ix + 1];
                const double qx = row[2 * (ix + s)], qy = row[2 * (ix + s) + 1];
                ar += px * qx + py * qy;
                ai += py * qx - px * qy;
                a0 += row0[ix] * row0[ix + s];
            }
        }
        acc[2 * s] = ar;
        acc[2 * s + 1] = ai;
        acc0[s] = a0;
    }
}


/* ---- fill jitter of analyze_tolerance (tolerance.cpp:20-31) ----------------------------
 * The reference draws with the C++ standard library's std::mt19937 and
 * std::normal_distribution<double> (libstdc++).  Restated from their published
 * definitions, not from any library source:
 *   MT19937      Matsumoto & Nishimura, "Mersenne Twister", ACM TOMACS 8(1), 1998: state
 *                of 624 words seeded by x_i = 1812433253 (x_{i-1} ^ (x_{i-1} >> 30)) + i,
 *                twist with matrix 0x9908b0df at offset 397, tempering (11, 7/0x9d2c5680,
 *                15/0xefc60000, 18).  C++ [rand.predef] fixes its 10000th output from the
 *                default seed 5489 at 4123659995 (checked in tests/test_oracle.py).
 *   canonical    C++ [rand.util.canonical] for 53 bits from a 32-bit engine: two draws
 *                a, b and (a + b 2^32) / 2^64, a result of 1 replaced by the largest
 *                double below 1.
 *   normal       Marsaglia & Bray's polar method (SIAM Review 6(3), 1964): draw
 *                x, y = 2u - 1 until 0 < r2 = x^2 + y^2 <= 1, m = sqrt(-2 ln(r2) / r2);
 *                the pair gives y m now and x m on the next call; value * sigma + mean. */
typedef struct {
    uint32_t mt[624];
    int idx;
    int have_saved;
    double saved;
} orc_rng;

static void mt_seed(orc_rng *g, uint32_t s) {
    g->mt[0] = s;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
    g->have_saved = 0;
}

static uint32_t mt_next(orc_rng *g) {
    if (g->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            const uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            g->mt[i] = g->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

static double canonical53(orc_rng *g) {
    const double a = (double)mt_next(g);
    const double b = (double)mt_next(g);
    const double r = 4294967296.0;
    double v = (a + b * r) / (r * r);
    if (v >= 1.0) v = nextafter(1.0, 0.0);
    return v;
}

static double normal_draw(orc_rng *g, double mean, double sigma) {
    double v;
    if (g->have_saved) {
        g->have_saved = 0;
        v = g->saved;
    } else {
        double x, y, r2;
        do {
            x = 2.0 * canonical53(g) - 1.0;
            y = 2.0 * canonical53(g) - 1.0;
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0.0);
        const double m = sqrt(-2.0 * log(r2) / r2);
        g->saved = x * m;
        g->have_saved = 1;
        v = y * m;
    }
    return v * sigma + mean;
}

/* The n-th (1-based) output of an MT19937 seeded with s (for the known-answer test). */
uint32_t orc_mt19937_nth(uint32_t s, int64_t nth) {
    orc_rng g;
    mt_seed(&g, s);
    uint32_t v = 0;
    for (int64_t i = 0; i < nth; ++i) v = mt_next(&g);
    return v;
}

/* One trial's jittered fill map: fill_out[c] = clamp(fill_in[c] + N(0, sigma_fill),
 * 0.05, 0.95) in map order, sigma_fill = (sigma_nm / 1000) / period_um, engine seeded
 * seed + trial * 7919 + unsigned(sigma_nm * 13) (32-bit unsigned arithmetic). */
void orc_tolerance_fills(uint32_t seed, int trial, double sigma_nm, double period_um,
                         const double *fill_in, int64_t n, double *fill_out) {
    orc_rng g;
    mt_seed(&g, seed + (uint32_t)trial * 7919u + (uint32_t)(sigma_nm * 13.0));
    const double sigma_fill = (sigma_nm / 1000.0) / period_um;
    for (int64_t c = 0; c < n; ++c) {
        const double v = fill_in[c] + normal_draw(&g, 0.0, sigma_fill);
        fill_out[c] = v < 0.05 ? 0.05 : (0.95 < v ? 0.95 : v);
    }
}
