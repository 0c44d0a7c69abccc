/* Body of the CPU restatement, instantiated once per number format by sht_ref.c with
 * REAL = float / double and SFX = _f32 / _f64.  TEST INFRASTRUCTURE (oracle), see sht_ref.c.
 */
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define CPX CAT(cpx, SFX)
#define FMA CAT(fma_, SFX)

typedef struct { REAL re, im; } CPX;

static inline REAL FMA(REAL a, REAL b, REAL c) { return (sizeof(REAL) == 4) ? fmaf((float)a, (float)b, (float)c) : fma(a, b, c); }

/* ---------------------------------------------------------------- mixed-radix FFT
 * Recursive decimation-in-time complex DFT  X[k] = sum_n x[n] w^{nk},  w = exp(sign 2 pi i / N).
 * N = p_0 p_1 ... ; radix 2, 3, 4, 5 butterflies are written out, any other factor uses a
 * plain DFT.  tw[k] = w^k for k < N (computed in double, rounded to REAL). */
typedef struct {
    int n, nfac;
    int fac[32];      /* radices, outermost first */
    CPX *tw;          /* n twiddles */
} CAT(fftplan, SFX);

static void CAT(fft_plan_init, SFX)(CAT(fftplan, SFX) *pl, int n, int sign) {
    pl->n = n;
    pl->nfac = 0;
    int r = n;
    while (r % 4 == 0) { pl->fac[pl->nfac++] = 4; r /= 4; }
    for (int p = 2; r > 1;) {
        if (r % p == 0) { pl->fac[pl->nfac++] = p; r /= p; }
        else p = (p == 2) ? 3 : p + 2;
    }
    pl->tw = (CPX *)malloc(sizeof(CPX) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        double a = sign * 2.0 * M_PI * (double)k / (double)n;
        pl->tw[k].re = (REAL)cos(a);
        pl->tw[k].im = (REAL)sin(a);
    }
}

static void CAT(fft_plan_free, SFX)(CAT(fftplan, SFX) *pl) { free(pl->tw); }

static inline CPX CAT(cmul, SFX)(CPX a, CPX b) {
    CPX c; c.re = a.re * b.re - a.im * b.im; c.im = a.re * b.im + a.im * b.re; return c;
}

/* out[0..N) = DFT of in[0], in[is], in[2 is], ...; fstride = twiddle stride (N_total / N). */
static void CAT(fft_rec, SFX)(const CAT(fftplan, SFX) *pl, CPX *out, const CPX *in, long is,
                              int fstride, const int *fac, int nfac, CPX *work) {
    int p = fac[0];
    int n = 1;
    for (int i = 0; i < nfac; ++i) n *= fac[i];
    int m = n / p;
    /* sub-DFTs of the p decimated subsequences, stored contiguously out[q m + u] */
    if (m == 1) {
        for (int q = 0; q < p; ++q) out[q] = in[q * is];
    } else {
        for (int q = 0; q < p; ++q)
            CAT(fft_rec, SFX)(pl, out + q * m, in + q * is, is * p, fstride * p, fac + 1, nfac - 1, work);
    }
    const CPX *tw = pl->tw;
    const int N = pl->n;
    for (int u = 0; u < m; ++u) {
        CPX t[64];
        CPX *tt = (p <= 64) ? t : work;
        for (int q = 0; q < p; ++q) {
            CPX v = out[q * m + u];
            tt[q] = (q == 0 || u == 0) ? v : CAT(cmul, SFX)(v, tw[(long)fstride * q * u % N]);
        }
        if (p == 2) {
            CPX a = tt[0], b = tt[1];
            out[u].re = a.re + b.re; out[u].im = a.im + b.im;
            out[u + m].re = a.re - b.re; out[u + m].im = a.im - b.im;
        } else if (p == 4) {
            /* w4 = tw[N/4] = +-i depending on sign */
            CPX w4 = tw[N / 4];
            CPX a0 = tt[0], a1 = tt[1], a2 = tt[2], a3 = tt[3];
            CPX s02 = {a0.re + a2.re, a0.im + a2.im}, d02 = {a0.re - a2.re, a0.im - a2.im};
            CPX s13 = {a1.re + a3.re, a1.im + a3.im}, d13 = {a1.re - a3.re, a1.im - a3.im};
            CPX r = CAT(cmul, SFX)(d13, w4);
            out[u].re = s02.re + s13.re; out[u].im = s02.im + s13.im;
            out[u + m].re = d02.re + r.re; out[u + m].im = d02.im + r.im;
            out[u + 2 * m].re = s02.re - s13.re; out[u + 2 * m].im = s02.im - s13.im;
            out[u + 3 * m].re = d02.re - r.re; out[u + 3 * m].im = d02.im - r.im;
        } else {
            /* plain DFT of length p using the global twiddles w_N^{(N/p) q s} */
            int step = N / p;
            for (int s = 0; s < p; ++s) {
                CPX acc = tt[0];
                for (int q = 1; q < p; ++q) {
                    CPX w = tw[(long)step * ((long)q * s % p)];
                    acc.re += tt[q].re * w.re - tt[q].im * w.im;
                    acc.im += tt[q].re * w.im + tt[q].im * w.re;
                }
                out[u + s * m] = acc;
            }
        }
    }
}

static void CAT(fft_exec, SFX)(const CAT(fftplan, SFX) *pl, CPX *out, const CPX *in, CPX *work) {
    CAT(fft_rec, SFX)(pl, out, in, 1, 1, pl->fac, pl->nfac, work);
}

/* -------------------------------------------------------------- spectral -> grid */
/* legendre.jl:2-33 (_fused_oddeven_matvec!) and :38-101 (_legendre!), then the per-ring
 * brfft of fourier.jl:57-77 done as one complex FFT of z = g_N + i g_S per ring. */
int CAT(oref_synthesis, SFX)(const oref_geom *g, const REAL *lam /* [j][lm] */,
                            const REAL *coeffs /* (K, nharm) complex interleaved */, int K,
                            int unscale_coslat, REAL *field /* (K, npoints) */, int nthreads) {
    const int n = g->nlat_half, lmax = g->lmax, nharm = g->nharm;
    int nfmax = 0;
    for (int j = 0; j < n; ++j) if (g->nlons[j] / 2 + 1 > nfmax) nfmax = g->nlons[j] / 2 + 1;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int err = 0;
#pragma omp parallel
    {
        CPX *gN = (CPX *)malloc(sizeof(CPX) * (size_t)nfmax * K);
        CPX *gS = (CPX *)malloc(sizeof(CPX) * (size_t)nfmax * K);
        int nlon_max = 2 * nfmax;
        CPX *z = (CPX *)malloc(sizeof(CPX) * (size_t)nlon_max);
        CPX *Z = (CPX *)malloc(sizeof(CPX) * (size_t)nlon_max);
        CPX *work = (CPX *)malloc(sizeof(CPX) * (size_t)nlon_max);
        if (!gN || !gS || !z || !Z || !work) err = 1;
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < n; ++j) {
            if (err) continue;
            const REAL *lj = lam + (size_t)j * nharm;
            const int nlon = g->nlons[j], nf = nlon / 2 + 1;
            memset(gN, 0, sizeof(CPX) * (size_t)nf * K);
            memset(gS, 0, sizeof(CPX) * (size_t)nf * K);
            long lm = 0;
            for (int m = 0; m <= g->mtrunc[j]; ++m) {
                const int L = lmax - m;
                const int lodd = L & 1, leven = L - lodd;
                const REAL *lv = lj + lm;
                for (int k = 0; k < K; ++k) {
                    const REAL *a = coeffs + 2 * ((size_t)k * nharm + lm);
                    REAL er = 0, ei = 0, orr = 0, oi = 0;
                    for (int l = 0; l < leven; l += 2) {
                        er = FMA(a[2 * l], lv[l], er);
                        ei = FMA(a[2 * l + 1], lv[l], ei);
                        orr = FMA(a[2 * l + 2], lv[l + 1], orr);
                        oi = FMA(a[2 * l + 3], lv[l + 1], oi);
                    }
                    REAL last = lodd ? lv[L - 1] : (REAL)0;
                    er = FMA(a[2 * (L - 1)], last, er);
                    ei = FMA(a[2 * (L - 1) + 1], last, ei);
                    CPX nn = {er + orr, ei + oi}, ss = {er - orr, ei - oi};
                    if (unscale_coslat) {
                        REAL c = (REAL)g->coslat_inv[j];
                        nn.re *= c; nn.im *= c; ss.re *= c; ss.im *= c;
                    }
                    gN[(size_t)k * nf + m] = nn;
                    gS[(size_t)k * nf + m] = ss;
                }
                lm += L;
            }
            /* brfft of north and south ring at once: Z[m] = gN[m] + i gS[m], Hermitian-extended,
             * Im of DC and Nyquist ignored */
            FFTPLAN_T pl;
            FFT_INIT(&pl, nlon, +1);
            const long rN = g->ring_start[j], rS = g->ring_start[2 * n - 1 - j];
            for (int k = 0; k < K; ++k) {
                const CPX *a = gN + (size_t)k * nf, *b = gS + (size_t)k * nf;
                for (int i = 0; i < nlon; ++i) { Z[i].re = 0; Z[i].im = 0; }
                Z[0].re = a[0].re; Z[0].im = b[0].re;
                for (int mm = 1; mm < nlon / 2; ++mm) {
                    Z[mm].re = a[mm].re - b[mm].im;  Z[mm].im = a[mm].im + b[mm].re;
                    Z[nlon - mm].re = a[mm].re + b[mm].im; Z[nlon - mm].im = b[mm].re - a[mm].im;
                }
                Z[nlon / 2].re = a[nlon / 2].re; Z[nlon / 2].im = b[nlon / 2].re;
                FFT_EXEC(&pl, z, Z, work);
                REAL *fk = field + (size_t)k * g->npoints;
                for (int i = 0; i < nlon; ++i) { fk[rN + i] = z[i].re; fk[rS + i] = z[i].im; }
            }
            FFT_FREE(&pl);
        }
        free(gN); free(gS); free(z); free(Z); free(work);
    }
    return err;
}

/* -------------------------------------------------------------- grid -> spectral */
/* fourier.jl:27-46,129-166 (rfft per ring), legendre.jl:104-126,131-198.  The reference
 * accumulates every coefficient over rings j = 1..nlat_half in ascending order; the loop
 * below keeps that order per coefficient and parallelises over orders m instead. */
int CAT(oref_analysis, SFX)(const oref_geom *g, const REAL *lam /* [j][lm] */,
                           const REAL *field /* (K, npoints) */, int K,
                           REAL *coeffs /* (K, nharm) complex */, int nthreads) {
    const int n = g->nlat_half, lmax = g->lmax, nharm = g->nharm;
    int nfmax = 0;
    for (int j = 0; j < n; ++j) if (g->nlons[j] / 2 + 1 > nfmax) nfmax = g->nlons[j] / 2 + 1;
    /* Fourier stage: per ring (parallel), f_north / f_south[j][k][m] */
    CPX *fN = (CPX *)malloc(sizeof(CPX) * (size_t)n * K * nfmax);
    CPX *fS = (CPX *)malloc(sizeof(CPX) * (size_t)n * K * nfmax);
    if (!fN || !fS) { free(fN); free(fS); return 1; }
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        int nlon_max = 2 * nfmax;
        CPX *z = (CPX *)malloc(sizeof(CPX) * (size_t)nlon_max);
        CPX *Z = (CPX *)malloc(sizeof(CPX) * (size_t)nlon_max);
        CPX *work = (CPX *)malloc(sizeof(CPX) * (size_t)nlon_max);
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < n; ++j) {
            const int nlon = g->nlons[j], nf = nlon / 2 + 1;
            FFTPLAN_T pl;
            FFT_INIT(&pl, nlon, -1);
            const long rN = g->ring_start[j], rS = g->ring_start[2 * n - 1 - j];
            for (int k = 0; k < K; ++k) {
                const REAL *fk = field + (size_t)k * g->npoints;
                for (int i = 0; i < nlon; ++i) { z[i].re = fk[rN + i]; z[i].im = fk[rS + i]; }
                FFT_EXEC(&pl, Z, z, work);
                CPX *a = fN + ((size_t)j * K + k) * nfmax, *b = fS + ((size_t)j * K + k) * nfmax;
                for (int mm = 0; mm < nf; ++mm) {
                    CPX p = Z[mm], q = Z[(nlon - mm) % nlon];
                    a[mm].re = (REAL)0.5 * (p.re + q.re); a[mm].im = (REAL)0.5 * (p.im - q.im);
                    b[mm].re = (REAL)0.5 * (p.im + q.im); b[mm].im = (REAL)0.5 * (q.re - p.re);
                }
            }
            FFT_FREE(&pl);
        }
        free(z); free(Z); free(work);
    }
    /* Legendre stage, parallel over orders m (each coefficient accumulated over j in order) */
#pragma omp parallel
    {
        CPX *ev = (CPX *)malloc(sizeof(CPX) * (size_t)K);
        CPX *od = (CPX *)malloc(sizeof(CPX) * (size_t)K);
#pragma omp for schedule(dynamic, 1)
        for (int m = 0; m < g->mmax; ++m) {
            const long off = g->offsets[m];
            const int L = lmax - m;
            const int lodd = L & 1, leven = L - lodd;
            for (int k = 0; k < K; ++k) {
                REAL *a = coeffs + 2 * ((size_t)k * nharm + off);
                for (int l = 0; l < 2 * L; ++l) a[l] = 0;
            }
            for (int j = 0; j < n; ++j) {
                if (m > g->mtrunc[j]) continue;
                const REAL dO = (REAL)g->solid_angles[j];
                for (int k = 0; k < K; ++k) {
                    CPX p = fN[((size_t)j * K + k) * nfmax + m], q = fS[((size_t)j * K + k) * nfmax + m];
                    ev[k].re = dO * (p.re + q.re); ev[k].im = dO * (p.im + q.im);
                    od[k].re = dO * (p.re - q.re); od[k].im = dO * (p.im - q.im);
                }
                const REAL *lv = lam + (size_t)j * nharm + off;
                for (int k = 0; k < K; ++k) {
                    REAL *a = coeffs + 2 * ((size_t)k * nharm + off);
                    for (int l = 0; l < leven; l += 2) {
                        a[2 * l] = FMA(lv[l], ev[k].re, a[2 * l]);
                        a[2 * l + 1] = FMA(lv[l], ev[k].im, a[2 * l + 1]);
                        a[2 * l + 2] = FMA(lv[l + 1], od[k].re, a[2 * l + 2]);
                        a[2 * l + 3] = FMA(lv[l + 1], od[k].im, a[2 * l + 3]);
                    }
                    REAL last = lodd ? lv[L - 1] : (REAL)0;
                    a[2 * (L - 1)] = FMA(last, ev[k].re, a[2 * (L - 1)]);
                    a[2 * (L - 1) + 1] = FMA(last, ev[k].im, a[2 * (L - 1) + 1]);
                }
            }
        }
        free(ev); free(od);
    }
    free(fN); free(fS);
    return 0;
}

#undef CPX
#undef FMA
