// Host side of libsht: plan construction (FP64 precompute), validation and the C ABI.
//
// The precompute follows SpectralTransform (SpeedyTransforms/src/spectral_transform.jl:
// 118-302): ring sizes and order (RingGrids/src/grids/octahedral_gaussian.jl:50-116,
// full_gaussian.jl:43-61), Gaussian latitudes and weights (full_gaussian.jl:48-50,
// quadrature_weights.jl:11,48-49), Legendre shortcut (legendre_shortcuts.jl:21-40),
// Legendre polynomials (spectral_transform.jl:167-172) computed here in FP64 and stored
// in the order the gfx950 kernels read them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/sht.h"
#include "sht_internal.h"

using sht::CallArgs;
using sht::DevPlan;
using sht::ColRef;
using sht::Tile;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

}  // namespace

void sht::set_last_error(const std::string &msg) { g_err = msg; }

namespace {

// ------------------------------------------------------------------ geometry

struct Geometry {
    int n = 0, nlat = 0, lmax = 0, mmax = 0, nharm = 0, nlon_max = 0;
    long long npoints = 0, fb_rows = 0;
    // mt: last order kept on ring j (the Legendre shortcut); mte: its running maximum over
    // rings 0..j, which decides the rings of an order in the Legendre stage (j0)
    std::vector<int> nlon, mt, mte, j0;
    std::vector<long long> ring_n, ring_s, fboff, lm_off;
    bool monotonic = true;
};

void gauss_legendre(int N, std::vector<double> &x, std::vector<double> &w);

int make_geometry(const sht_desc *d, Geometry &g) {
    if (!d) return fail(SHT_ERR_ARG, "null descriptor");
    if (d->grid != SHT_GRID_OCTAHEDRAL_GAUSSIAN && d->grid != SHT_GRID_FULL_GAUSSIAN)
        return fail(SHT_ERR_UNSUPPORTED, "grid kind must be octahedral or full Gaussian");
    if (d->nlat_half < 1 || d->nlat_half > 4096) return fail(SHT_ERR_ARG, "nlat_half out of range");
    if (d->mmax < 1 || d->lmax < d->mmax || d->lmax > 8192)
        return fail(SHT_ERR_ARG, "need 1 <= mmax <= lmax <= 8192");
    if (d->nlayers < 1) return fail(SHT_ERR_ARG, "nlayers must be >= 1");
    if (d->shortcut < SHT_SHORTCUT_LINEAR || d->shortcut > SHT_SHORTCUT_LINCUB_COSLAT)
        return fail(SHT_ERR_UNSUPPORTED, "unknown Legendre shortcut");
    if (d->precision < SHT_PREC_FP32 || d->precision > SHT_PREC_BF16X3)
        return fail(SHT_ERR_ARG, "unknown precision mode");
    const int n = d->nlat_half;
    g.n = n;
    g.nlat = 2 * n;
    g.lmax = d->lmax;
    g.mmax = d->mmax;
    g.nharm = d->lmax * d->mmax - d->mmax * (d->mmax - 1) / 2;
    g.nlon.resize(n);
    for (int j = 0; j < n; ++j)
        g.nlon[j] = d->grid == SHT_GRID_OCTAHEDRAL_GAUSSIAN ? 16 + 4 * (j + 1) : 4 * n;
    g.nlon_max = *std::max_element(g.nlon.begin(), g.nlon.end());
    // ring ranges: northern rings 0..n-1, then southern rings mirrored
    g.ring_n.resize(n);
    g.ring_s.resize(n);
    long long pos = 0;
    for (int j = 0; j < n; ++j) { g.ring_n[j] = pos; pos += g.nlon[j]; }
    for (int j = n - 1; j >= 0; --j) { g.ring_s[j] = pos; pos += g.nlon[j]; }
    g.npoints = pos;
    // Legendre shortcut (legendre_shortcuts.jl:21-49; spectral_transform.jl:154-155):
    // mt_j = min(shortcut(nlon_j, latd_j), mmax - 1), with latd_j = acosd(x_j) - 90 of the
    // ascending Gauss nodes x_j (full_gaussian.jl:48-50) for the latitude-dependent ones
    g.mt.resize(n);
    std::vector<double> xg, wg;
    if (d->shortcut >= SHT_SHORTCUT_LINQUAD_COSLAT2) gauss_legendre(2 * n, xg, wg);
    for (int j = 0; j < n; ++j) {
        int s;
        switch (d->shortcut) {
            case SHT_SHORTCUT_LINEAR: s = g.nlon[j] / 2; break;
            case SHT_SHORTCUT_QUADRATIC: s = g.nlon[j] / 3; break;
            case SHT_SHORTCUT_CUBIC: s = g.nlon[j] / 4; break;
            default: {
                const double latd = std::acos(xg[j]) * (180.0 / M_PI) - 90.0;  // ascending nodes: north first
                const double c = std::cos(latd * (M_PI / 180.0));
                const double den = d->shortcut == SHT_SHORTCUT_LINQUAD_COSLAT2 ? 2.0 + c * c : 2.0 + 2.0 * c;
                s = (int)std::floor((double)g.nlon[j] / den);
            }
        }
        g.mt[j] = std::min(s, g.mmax - 1);
    }
    g.mte.resize(n);
    for (int j = 0; j < n; ++j) {
        g.mte[j] = j == 0 ? g.mt[0] : std::max(g.mte[j - 1], g.mt[j]);
        if (g.mte[j] != g.mt[j]) g.monotonic = false;
        // the Legendre stage writes orders up to mte_j into ring j's Fourier rows
        if (g.mte[j] > g.nlon[j] / 2) return fail(SHT_ERR_UNSUPPORTED, "shortcut keeps more orders than a ring holds");
    }
    g.fboff.resize(n);
    long long rows = 0;
    for (int j = 0; j < n; ++j) { g.fboff[j] = rows; rows += g.nlon[j] / 2 + 1; }
    g.fb_rows = rows;
    g.j0.resize(g.mmax);
    for (int m = 0; m < g.mmax; ++m) {
        int j = 0;
        while (j < n && g.mte[j] < m) ++j;
        g.j0[m] = j;
    }
    g.lm_off.resize(g.mmax);
    for (int m = 0; m < g.mmax; ++m) g.lm_off[m] = (long long)m * g.lmax - (long long)m * (m - 1) / 2;
    return SHT_OK;
}

// ------------------------------------------------------------- quadrature, Λ

void legendre_pn(int N, double x, double &p, double &dp) {
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= N; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    p = N == 0 ? 1.0 : p1;
    dp = N * (x * p1 - p0) / (x * x - 1.0);
}

// Gauss-Legendre nodes (ascending) and weights: Newton on P_N from the Tricomi-type
// initial guess, weights 2 / ((1-x^2) P_N'(x)^2).
void gauss_legendre(int N, std::vector<double> &x, std::vector<double> &w) {
    x.resize(N);
    w.resize(N);
    for (int i = 0; i < (N + 1) / 2; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (N + 0.5));  // i-th largest root
        for (int it = 0; it < 100; ++it) {
            double p, dp;
            legendre_pn(N, z, p, dp);
            const double dz = p / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-15) break;
        }
        double p, dp;
        legendre_pn(N, z, p, dp);
        z -= p / dp;  // one more step at full precision
        legendre_pn(N, z, p, dp);
        const double wi = 2.0 / ((1.0 - z * z) * dp * dp);
        x[N - 1 - i] = z;
        x[i] = -z;
        w[i] = w[N - 1 - i] = wi;
    }
    if (N % 2 == 1) x[N / 2] = 0.0;
}

// λ_l^m(x) for one order m and l = m..lmax-1, from λ_m^m (given) by the three-term
// recursion of the orthonormal associated Legendre functions.
void legendre_order(int lmax, int m, double x, double pmm, double *out) {
    out[0] = pmm;
    if (m + 1 >= lmax) return;
    double p0 = pmm, p1 = std::sqrt(2.0 * m + 3.0) * x * pmm;
    out[1] = p1;
    for (int l = m + 2; l < lmax; ++l) {
        const double a = std::sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
        const double b = std::sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
        const double p2 = a * (x * p1 - b * p0);
        p0 = p1;
        p1 = p2;
        out[l - m] = p2;
    }
}

// λ_m^m from λ_{m-1}^{m-1}: Condon-Shortley phase, λ_0^0 = 1/sqrt(4 pi)
inline double next_pmm(double pmm_prev, int m, double s) {
    return m == 0 ? 1.0 / std::sqrt(4.0 * M_PI) : -std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * s * pmm_prev;
}

// ---------------------------------------------------------- 16-bit conversions

uint16_t to_half(double v) {
    const _Float16 h = (_Float16)(float)v;
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}
double half_value(uint16_t u) {
    _Float16 h;
    std::memcpy(&h, &u, 2);
    return (double)(float)h;
}
uint16_t to_bf16(double v) {
    const float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
double bf16_value(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

// power of two mapping amax into [2^14, 2^15)
double fp16_scale(double amax) {
    if (!(amax > 0.0)) return 1.0;
    int e;
    std::frexp(amax, &e);
    int sh = std::max(-120, std::min(120, 15 - e));
    return std::ldexp(1.0, sh);
}

// K position within a synthesis chunk of the chunk's l' = ci (ci < 32): the staging thread g
// of a column holds l' = g + 8i (i = 0..3, degree pairs read as 128-byte runs) and writes them
// to positions 4g + i, so the Λ fragments take the same order.
inline int syn_kpos(int ci) { return 4 * (ci % 8) + ci / 8; }

template <class T>
int dev_upload(std::vector<void *> &allocs, const std::vector<T> &h, const T **out) {
    *out = nullptr;
    if (h.empty()) return SHT_OK;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, h.size() * sizeof(T));
    if (e != hipSuccess) return fail(SHT_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(p);
    e = hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(SHT_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    *out = static_cast<const T *>(p);
    return SHT_OK;
}

}  // namespace

// Optional per-kernel timing: events recorded around every launch on the call's stream.
struct Profile {
    bool on = false;
    int max_calls = 0;
    int ncalls[SHT_NKERNELS] = {0};
    std::vector<hipEvent_t> ev;  // [(kernel * max_calls + call) * 2 + {start, stop}]
};

// Ring FFT factorisation policy (a tuning knob of the autotuner).
struct FFTPolicy {
    int mfma_min = 17;       // a largest prime >= this runs first, as the MFMA pass
    int eights = 1;          // powers of two as 4s (0), 8s (1) or 16s (2), then the rest as 8, 4, 2
    bool odd_first = false;  // small odd radices before the powers of two
};
// Every policy's FFT tables live on the device from plan creation on, so choosing one
// (autotuning, pinning) changes host state only and never rewrites memory that kernels in
// flight on other streams may read.
constexpr int kPolicies = 12;
inline int policy_index(const FFTPolicy &p) { return (p.mfma_min == 11 ? 6 : 0) + 2 * p.eights + (p.odd_first ? 1 : 0); }
inline FFTPolicy policy_of(int i) {
    FFTPolicy p;
    p.mfma_min = i >= 6 ? 11 : 17;
    p.eights = (i % 6) / 2;
    p.odd_first = (i & 1) != 0;
    return p;
}
// Streams one plan serves with their own Fourier scratch (each at most one buffer of the
// plan's capacity; more streams is an error, not an eviction of scratch in use).
constexpr int kMaxPlanStreams = 16;

// A launch order for workgroups grouped in units that share data (tiles of one unit run
// back to back on one XCD): units, heaviest first, go greedily to the least-loaded XCD;
// block b then takes slot b / 8 of XCD b % 8's list.  Lists are padded to equal length
// with padding tiles (m = -1).
struct LaunchUnit {
    double cost;
    std::vector<Tile> tiles;
};

// Light-first rotation of a heaviest-first list for `slots` resident workgroups: the tiles
// beyond the slots (the lightest) move to the front, so that they finish early and the
// tiles that start late are light ones starting early rather than tiles that wait for the
// first heavy ones to end (which left a tail of few resident workgroups).  slots = 0: as is.
template <class T>
static void light_first(std::vector<T> &list, long long slots) {
    if (slots <= 0 || (long long)list.size() <= slots) return;
    const long long e = (long long)list.size() - slots;
    std::rotate(list.begin(), list.end() - e, list.end());
}

static std::vector<Tile> xcd_order(std::vector<LaunchUnit> units, long long slots_per_xcd = 0) {
    std::stable_sort(units.begin(), units.end(),
                     [](const LaunchUnit &a, const LaunchUnit &b) { return a.cost > b.cost; });
    std::vector<std::vector<Tile>> lists(sht::kXCDs);
    std::vector<double> load(sht::kXCDs, 0.0);
    for (const LaunchUnit &u : units) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[x] += u.cost;
        lists[x].insert(lists[x].end(), u.tiles.begin(), u.tiles.end());
    }
    for (auto &l : lists) light_first(l, slots_per_xcd);
    size_t len = 0;
    for (const auto &l : lists) len = std::max(len, l.size());
    std::vector<Tile> out(len * sht::kXCDs, Tile{0, -1, 0, 0});
    for (int x = 0; x < sht::kXCDs; ++x)
        for (size_t q = 0; q < lists[x].size(); ++q) out[q * sht::kXCDs + x] = lists[x][q];
    return out;
}

// Legendre units: the tiles of one order m (all ring or degree groups, all column groups).
static std::vector<LaunchUnit> legendre_units(const std::vector<Tile> &t, const std::vector<double> &c, int call,
                                              int cgs) {
    std::vector<LaunchUnit> units;
    std::vector<int> unit_of_m;
    for (size_t q = 0; q < t.size(); ++q) {
        const int m = t[q].m;
        if (m >= (int)unit_of_m.size()) unit_of_m.resize(m + 1, -1);
        if (unit_of_m[m] < 0) {
            unit_of_m[m] = (int)units.size();
            units.push_back({0.0, {}});
        }
        LaunchUnit &u = units[unit_of_m[m]];
        for (int cg = 0; cg < cgs; ++cg) {
            u.tiles.push_back({call, m, t[q].blk, cg});
            u.cost += c[q];
        }
    }
    return units;
}

struct sht_plan {
    sht_desc d{};
    Geometry g;
    DevPlan P{};                  // host copy of the device view under the chosen FFT policy
    const DevPlan *dP = nullptr;  // its device copy (an element of dPv)
    DevPlan Pv[kPolicies]{};      // device views per FFT policy (same tables but the FFT's)
    const DevPlan *dPv = nullptr;
    bool pol_ok[kPolicies] = {};  // the policy factorises every ring
    int pol = 0;                  // chosen policy
    std::vector<void *> allocs;
    // single-call Legendre launch orders, one per column-group count cgs = 1..max, grouped
    // by order m per XCD ([1]) or plain heavy-first ([0])
    // Legendre launch orders [variant][column groups - 1]: variant bit 0 groups the tiles of
    // an order on one XCD, bit 1 rotates the lightest tiles beyond the resident slots first
    std::vector<const Tile *> syn_tiles[4], ana_tiles[4];
    std::vector<int> nsyn[4], nana[4];
    // tuned choices (sht_plan_autotune): FFT factorisation, Legendre order grouping, FFT
    // synthesis column-group pairing, Fourier kernel kind and its columns per workgroup
    int leg_xcd = 1, fft_pair = 1;
    int fourier = sht::kFourierRadix, dft_cg = 4;
    int tuned = 0;
    // dynamic LDS bytes of the DFT kernels: [synthesis][FP16][8 columns]
    int dft_smem[2][2][2] = {};
    // host copies for batches: Legendre tiles (heavy first) with their costs, FFT ring order
    // and cost per policy
    std::vector<Tile> h_syn, h_ana;
    std::vector<double> h_syn_cost, h_ana_cost, h_ring_cost_v[kPolicies];
    std::vector<int> h_ring_order_v[kPolicies];
    // Scratch per stream: the Fourier buffer (fb_bytes), or on matrix plans also the dense
    // GEMMs' split-K sums (dense_bytes; a call uses one or the other); created on a stream's
    // first call, freed with the plan.
    size_t fb_bytes = 0, dense_bytes = 0;
    std::mutex scratch_mutex;
    std::vector<std::pair<void *, void *>> scratch;  // (stream, buffer)
    long long device_bytes = 0;
    Profile prof;
    // dense matrix transform (sht_plan_create_matrix): sht_synthesis / sht_analysis run as
    // one FP32 GEMM each with these column-major matrices (complex sides interleaved re, im)
    bool matrix = false;
    float *mat_syn = nullptr;      // npoints x 2 nharm (fragment order): columns (Re B, -Im B) of field = Re(B a)
    float *mat_ana = nullptr;      // 2 nharm x npoints (fragment order): rows (Re F, Im F) of a = F f
    float *pt_coslat_inv = nullptr;  // 1/cos(lat) per grid point (unscale_coslat)
    // Graph replay of single calls (sht_plan_set_graphs; the reference's per-transform graph
    // cache, SpeedyTransforms/ext/gpu_graphs_common.jl:23,121,293-331): an explicitly built,
    // instantiated graph of a call's two stage kernels per distinct call (direction, buffers,
    // strides, K, flags, scratch, tuning), at most kMaxGraphs per direction; a null entry
    // marks a call whose graph failed, which then launches directly.
    bool graphs = false, graph_force_failure = false;
    std::mutex graph_mutex;
    std::map<std::string, hipGraphExec_t> graph_cache[2];
    int graph_failed = 0;
};

namespace {
// Brackets one kernel launch with profiling events when profiling is on.
struct ProfScope {
    sht_plan *pl;
    int k;
    hipStream_t s;
    bool rec = false;
    ProfScope(sht_plan *p, int kernel, void *stream) : pl(p), k(kernel), s(reinterpret_cast<hipStream_t>(stream)) {
        Profile &pr = pl->prof;
        if (pr.on && pr.ncalls[k] < pr.max_calls) {
            rec = true;
            (void)hipEventRecord(pr.ev[((size_t)k * pr.max_calls + pr.ncalls[k]) * 2], s);
        }
    }
    ~ProfScope() {
        Profile &pr = pl->prof;
        if (rec) {
            (void)hipEventRecord(pr.ev[((size_t)k * pr.max_calls + pr.ncalls[k]) * 2 + 1], s);
            ++pr.ncalls[k];
        }
    }
};
}  // namespace

namespace {

struct DeviceGuard {
    int prev = -1;
    bool changed = false;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) {
            changed = hipSetDevice(dev) == hipSuccess;
        }
    }
    ~DeviceGuard() {
        if (changed) (void)hipSetDevice(prev);
    }
};

void free_plan(sht_plan *p) {
    if (!p) return;
    for (auto &gc : p->graph_cache)
        for (auto &kv : gc)
            if (kv.second) (void)hipGraphExecDestroy(kv.second);
    for (hipEvent_t e : p->prof.ev) (void)hipEventDestroy(e);
    for (void *a : p->allocs) (void)hipFree(a);
    for (auto &sb : p->scratch) (void)hipFree(sb.second);
    delete p;
}

// The scratch of `stream` on this plan (allocated on the stream's first call), or null
// with the error recorded.  Calls on one stream are ordered, so they share it safely;
// calls on different streams never share one.
void *stream_scratch(sht_plan *p, void *stream) {
    std::lock_guard<std::mutex> lock(p->scratch_mutex);
    for (auto &sb : p->scratch)
        if (sb.first == stream) return sb.second;
    if ((int)p->scratch.size() >= kMaxPlanStreams) {
        fail(SHT_ERR_ALLOC, "a plan serves at most " + std::to_string(kMaxPlanStreams) +
                                " streams with their own scratch");
        return nullptr;
    }
    void *buf = nullptr;
    const hipError_t e = hipMalloc(&buf, std::max<size_t>(std::max(p->fb_bytes, p->dense_bytes), 256));
    if (e != hipSuccess) {
        fail(SHT_ERR_ALLOC, std::string("hipMalloc stream scratch: ") + hipGetErrorString(e));
        return nullptr;
    }
    p->scratch.push_back({stream, buf});
    return buf;
}

// Radices of the ring FFT, in pass order: the largest prime >= policy.mfma_min first (the
// first pass runs as an MFMA GEMM and needs no twiddles), then the powers of two (8s or
// 4s) and 2, 3, 5, 7, 11, 13 (odd ones first or last), then any further prime >= 11.
// Returns the count, or -1 if a radix has no kernel.
int factorize(int N, int *fac, const FFTPolicy &pol) {
    int nf = 0, r = N;
    std::vector<int> big;  // primes >= 11, largest first
    int tmp = r;
    for (int q : {2, 3, 5, 7})
        while (tmp % q == 0) tmp /= q;
    for (int p = 11; p <= tmp; p += 2) {  // tmp has no factor below 11: p divides only if prime
        while (tmp % p == 0) {
            big.push_back(p);
            tmp /= p;
        }
    }
    std::sort(big.begin(), big.end(), [](int a, int b) { return a > b; });
    size_t used = 0;
    if (!big.empty() && big[0] >= std::min(pol.mfma_min, sht::kMfmaMinRadix)) {
        if (big[0] > sht::kMaxMfmaRadix) return -1;
        for (int ncol = 1; ncol <= sht::kFFTCols; ++ncol)
            if (sht::prime_rows_per_wave(big[0], N / big[0], ncol) > sht::kPrimeRowTiles) return -1;
        fac[nf++] = big[0];
        r /= big[0];
        used = 1;
    }
    auto pow2 = [&]() {
        if (pol.eights == 2) {
            while (r % 16 == 0) { fac[nf++] = 16; r /= 16; }
        }
        if (pol.eights >= 1) {
            while (r % 8 == 0) { fac[nf++] = 8; r /= 8; }
        }
        while (r % 4 == 0) { fac[nf++] = 4; r /= 4; }
        while (r % 2 == 0) { fac[nf++] = 2; r /= 2; }
    };
    auto odd = [&]() {
        for (int q : {3, 5, 7})
            while (r % q == 0) { fac[nf++] = q; r /= q; }
    };
    if (pol.odd_first) {
        odd();
        pow2();
    } else {
        pow2();
        odd();
    }
    for (size_t i = used; i < big.size(); ++i) {
        if (big[i] > sht::kMaxRegisterRadix) return -1;
        fac[nf++] = big[i];
        r /= big[i];
    }
    return r == 1 ? nf : -1;
}

// Per-ring pass descriptors of a factorisation (radix, l, M and their division multipliers).
std::vector<sht::FFTPass> pass_table(const Geometry &g, const std::vector<int> &fac, const std::vector<int> &nfac) {
    std::vector<sht::FFTPass> t((size_t)g.n * sht::kMaxFactors, sht::FFTPass{1, 1, 1, 0u, 0u, {0, 0, 0}});
    for (int j = 0; j < g.n; ++j) {
        int l = 1;
        for (int f = 0; f < nfac[j]; ++f) {
            const int R = fac[(size_t)j * sht::kMaxFactors + f], M = g.nlon[j] / (l * R);
            t[(size_t)j * sht::kMaxFactors + f] = sht::FFTPass{R, l, M, sht::fastdiv_mul(M), sht::fastdiv_mul(R), {0, 0, 0}};
            l *= R;
        }
    }
    return t;
}

// FFT tables of a policy: radices per ring, modelled cost and heavy-first ring order.
int fft_tables(const Geometry &g, const FFTPolicy &pol, std::vector<int> &fac, std::vector<int> &nfac,
               std::vector<double> &cost, std::vector<int> &order) {
    const int n = g.n;
    fac.assign((size_t)n * sht::kMaxFactors, 1);
    nfac.assign(n, 0);
    cost.assign(n, 0.0);
    for (int j = 0; j < n; ++j) {
        nfac[j] = factorize(g.nlon[j], &fac[(size_t)j * sht::kMaxFactors], pol);
        if (nfac[j] < 0)
            return fail(SHT_ERR_UNSUPPORTED,
                        "ring length " + std::to_string(g.nlon[j]) + " has a prime factor the FFT cannot run");
        double c = 0;
        for (int f = 0; f < nfac[j]; ++f) {
            const int r = fac[(size_t)j * sht::kMaxFactors + f];
            c += r <= 16 ? 1.0 : (f == 0 ? 0.25 * r : 0.5 * r);  // MFMA first pass or register butterflies
        }
        cost[j] = c * g.nlon[j];
    }
    order.resize(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    return SHT_OK;
}

// Tables of the MFMA prime pass (see DevPlan::prime_tab) for every radix r >= 11 that is
// the first pass of some ring under some factorisation policy.
void prime_tables(const std::vector<int> (&fac_v)[kPolicies], const std::vector<int> (&nfac_v)[kPolicies],
                  const bool (&ok)[kPolicies], int n, std::vector<uint16_t> &tab, std::vector<long long> &off) {
    std::vector<bool> used(sht::kMaxMfmaRadix + 1, false);
    for (int q = 0; q < kPolicies; ++q) {
        if (!ok[q]) continue;
        for (int j = 0; j < n; ++j) {
            if (nfac_v[q][j] < 1) continue;
            const int r = fac_v[q][(size_t)j * sht::kMaxFactors];
            if (r >= sht::kMfmaMinPrime && r <= sht::kMaxMfmaRadix && (r & 1)) used[r] = true;
        }
    }
    off.assign(sht::kMaxMfmaRadix + 1, 0);
    tab.clear();
    for (int r = sht::kMfmaMinPrime; r <= sht::kMaxMfmaRadix; ++r) {
        if (!used[r]) continue;
        const int h = (r - 1) / 2, Mt = (h + 16) / 16, KC = (h + 31) / 32;
        off[r] = (long long)(tab.size() / 8);
        tab.resize(tab.size() + (size_t)Mt * KC * 64 * 6 * 8, 0);
        for (int mt = 0; mt < Mt; ++mt)
            for (int kc = 0; kc < KC; ++kc)
                for (int lane = 0; lane < 64; ++lane)
                    for (int e = 0; e < 8; ++e) {
                        const int s = 16 * mt + lane % 16, t = 32 * kc + 8 * (lane / 16) + e + 1;
                        double v[2] = {0.0, 0.0};
                        if (s <= h && t <= h) {
                            const double a = 2.0 * M_PI * (double)(((long long)s * t) % r) / (double)r;
                            v[0] = std::cos(a);
                            v[1] = std::sin(a);
                        }
                        const size_t base = (size_t)8 * (off[r] + ((size_t)(mt * KC + kc) * 64 + lane) * 6) + e;
                        for (int cs = 0; cs < 2; ++cs) {
                            const uint16_t hi = to_bf16(v[cs]);
                            const double r1 = v[cs] - bf16_value(hi);
                            const uint16_t mid = to_bf16(r1);
                            const uint16_t lo = to_bf16(r1 - bf16_value(mid));
                            tab[base + (size_t)(3 * cs + 0) * 8] = hi;
                            tab[base + (size_t)(3 * cs + 1) * 8] = mid;
                            tab[base + (size_t)(3 * cs + 2) * 8] = lo;
                        }
                    }
    }
    if (tab.empty()) tab.assign(8, 0);
}

// Matrices of the ring DFT on MFMA (see DevPlan::dft32 / dft16): for every ring of
// N = 4p points, cos and sin of 2 pi r t / p for rows r = 0..p/2 and t = 1..(p-1)/2 in FP64,
// zero outside, stored FP32 and (split plans) as FP16 hi / lo of 2^14 times the value.
void dft_tables(const Geometry &g, bool split, std::vector<float> &t32, std::vector<long long> &o32,
                std::vector<uint16_t> &t16, std::vector<long long> &o16) {
    o32.assign(g.n, 0);
    o16.assign(g.n, 0);
    for (int j = 0; j < g.n; ++j) {
        const int p = g.nlon[j] / 4, h = (p - 1) / 2, R = p / 2 + 1, Rt = (R + 15) / 16;
        auto cs = [&](int r, int t, double &c, double &sn) {
            c = sn = 0.0;
            if (r >= R || t < 1 || t > h) return;
            const double a = 2.0 * M_PI * (double)(((long long)r * t) % p) / (double)p;
            c = std::cos(a);
            sn = std::sin(a);
        };
        const int KC = (h + sht::kDftChunk32 - 1) / sht::kDftChunk32;
        o32[j] = (long long)t32.size();
        t32.resize(t32.size() + (size_t)Rt * KC * 64 * 8, 0.f);
        for (int rt = 0; rt < Rt; ++rt)
            for (int ch = 0; ch < KC; ++ch)
                for (int lane = 0; lane < 64; ++lane)
                    for (int st = 0; st < 4; ++st) {
                        double c, sn;
                        cs(16 * rt + lane % 16, sht::kDftChunk32 * ch + 4 * (lane / 16) + st + 1, c, sn);
                        const size_t idx = (size_t)o32[j] + (((size_t)rt * KC + ch) * 64 + lane) * 8;
                        t32[idx + st] = (float)c;
                        t32[idx + 4 + st] = (float)sn;
                    }
        if (!split) continue;
        const int KH = (h + sht::kDftChunk16 - 1) / sht::kDftChunk16;
        o16[j] = (long long)t16.size();
        t16.resize(t16.size() + (size_t)Rt * KH * 64 * 32, 0);
        for (int rt = 0; rt < Rt; ++rt)
            for (int ch = 0; ch < KH; ++ch)
                for (int lane = 0; lane < 64; ++lane)
                    for (int e = 0; e < 8; ++e) {
                        double c, sn;
                        cs(16 * rt + lane % 16, sht::kDftChunk16 * ch + 8 * (lane / 16) + e + 1, c, sn);
                        c *= sht::kDftAScale;
                        sn *= sht::kDftAScale;
                        const size_t idx = (size_t)o16[j] + (((size_t)rt * KH + ch) * 64 + lane) * 32;
                        t16[idx + e] = to_half(c);
                        t16[idx + 8 + e] = to_half(c - half_value(t16[idx + e]));
                        t16[idx + 16 + e] = to_half(sn);
                        t16[idx + 24 + e] = to_half(sn - half_value(t16[idx + 16 + e]));
                    }
    }
}

int build_plan(sht_plan *pl) {
    const sht_desc &d = pl->d;
    long long bytes_dft = 0;
    Geometry &g = pl->g;
    const int n = g.n, lmax = g.lmax, mmax = g.mmax;
    const int mode = d.precision;
    const int adeg = sht::ana_degs_per_wg(mode);  // degrees per analysis workgroup
    std::vector<void *> &A = pl->allocs;
    DevPlan &P = pl->P;
    int rc;

    // latitudes: nodes ascending, ring j (north first) has cos(colat) = -x_j
    std::vector<double> x, w;
    gauss_legendre(2 * n, x, w);
    std::vector<double> cosc(n), sinc(n), coslat_inv(n), dOmega(n);
    for (int j = 0; j < n; ++j) {
        cosc[j] = -x[j];
        sinc[j] = std::sqrt(std::max(0.0, 1.0 - x[j] * x[j]));
        coslat_inv[j] = 1.0 / sinc[j];  // cos(lat) = sin(colat)
        dOmega[j] = w[j] * 2.0 * M_PI / g.nlon[j];
    }

    // ---------------------------------------------------- Λ in kernel order
    std::vector<long long> syn_off(mmax), ana_off(mmax), syn_soff(mmax), ana_soff(mmax);
    long long syn_total = 0, ana_total = 0, syn_stotal = 0, ana_stotal = 0;
    const bool split = mode != SHT_PREC_FP32;
    for (int m = 0; m < mmax; ++m) {
        const int J = n - g.j0[m], L = lmax - m, Le = (L + 1) / 2;
        const int Jt = (J + 15) / 16, nch = (Le + sht::kChunk - 1) / sht::kChunk;
        const int NG = (L + adeg - 1) / adeg, nt = NG * (adeg / 32);
        const int JC = (J + sht::kChunk - 1) / sht::kChunk;
        syn_off[m] = syn_total;
        ana_off[m] = ana_total;
        syn_soff[m] = syn_stotal;
        ana_soff[m] = ana_stotal;
        syn_total += (long long)2 * Jt * nch * sht::kChunk * 16;  // same count in both layouts
        ana_total += (long long)2 * nt * JC * sht::kChunk * 16;
        syn_stotal += (long long)Jt * 16;
        ana_stotal += (long long)2 * nt * 16;
    }
    std::vector<float> syn_f32, ana_f32, syn_sinv, ana_sinv;
    std::vector<uint16_t> syn_hi, syn_lo, syn_mid, ana_hi, ana_lo, ana_mid;
    const bool three = mode == SHT_PREC_BF16X3;  // BF16: hi + mid + lo
    if (split) {
        syn_hi.assign(syn_total, 0); syn_lo.assign(syn_total, 0);
        ana_hi.assign(ana_total, 0); ana_lo.assign(ana_total, 0);
        if (three) {
            syn_mid.assign(syn_total, 0);
            ana_mid.assign(ana_total, 0);
        }
        syn_sinv.assign(syn_stotal, 1.f); ana_sinv.assign(ana_stotal, 1.f);
    } else {
        syn_f32.assign(syn_total, 0.f);
        ana_f32.assign(ana_total, 0.f);
    }
    // FP16: v s = hi + lo; BF16: v = hi + mid + lo (each part rounds the FP64 remainder)
    auto split_store = [&](double v, double s, std::vector<uint16_t> &hi, std::vector<uint16_t> &mid,
                           std::vector<uint16_t> &lo, long long idx) {
        if (mode == SHT_PREC_FP16X3) {
            const double xs = v * s;
            hi[idx] = to_half(xs);
            lo[idx] = to_half(xs - half_value(hi[idx]));
        } else {
            hi[idx] = to_bf16(v);
            const double r = v - bf16_value(hi[idx]);
            mid[idx] = to_bf16(r);
            lo[idx] = to_bf16(r - bf16_value(mid[idx]));
        }
    };

    std::vector<double> pmm(n, 0.0);
    std::vector<double> lam;  // [j - j0][l - m]
    for (int m = 0; m < mmax; ++m) {
        for (int j = 0; j < n; ++j) pmm[j] = next_pmm(pmm[j], m, sinc[j]);
        const int j0 = g.j0[m], J = n - j0, L = lmax - m, Le = (L + 1) / 2;
        if (J == 0) continue;
        lam.assign((size_t)J * L, 0.0);
        for (int jj = 0; jj < J; ++jj) legendre_order(lmax, m, cosc[j0 + jj], pmm[j0 + jj], &lam[(size_t)jj * L]);
        auto LAM = [&](int l, int j) -> double { return lam[(size_t)(j - j0) * L + (l - m)]; };
        const int Jt = (J + 15) / 16, nch = (Le + sht::kChunk - 1) / sht::kChunk;
        const int NG = (L + adeg - 1) / adeg, nt = NG * (adeg / 32);
        const int JC = (J + sht::kChunk - 1) / sht::kChunk;
        // synthesis: rows = rings, K = l' per parity
        std::vector<double> srow_scale(Jt * 16, 1.0);
        if (mode == SHT_PREC_FP16X3) {
            for (int t = 0; t < Jt; ++t)
                for (int r = 0; r < 16; ++r) {
                    const int j = j0 + 16 * t + r;
                    if (j >= n) continue;
                    double amax = 0.0;
                    for (int l = m; l < lmax; ++l) amax = std::max(amax, std::fabs(LAM(l, j)));
                    const double s = fp16_scale(amax);
                    srow_scale[t * 16 + r] = s;
                    syn_sinv[syn_soff[m] + t * 16 + r] = (float)(1.0 / s);
                }
        }
        for (int p = 0; p < 2; ++p)
            for (int t = 0; t < Jt; ++t)
                for (int lp = 0; lp < nch * sht::kChunk; ++lp) {
                    const int l = m + 2 * lp + p;
                    if (l >= lmax) continue;
                    const int c = lp / sht::kChunk, kq = syn_kpos(lp % sht::kChunk);
                    for (int r = 0; r < 16; ++r) {
                        const int j = j0 + 16 * t + r;
                        if (j >= n) continue;
                        const double v = LAM(l, j);
                        if (!split) {
                            syn_f32[syn_off[m] + ((long long)(p * Jt + t) * (nch * sht::kChunk) + c * sht::kChunk + kq) * 16 + r] =
                                (float)v;
                        } else {
                            const int grp = kq / 8, e = kq % 8, lane = grp * 16 + r;
                            const long long idx = syn_off[m] + (((long long)(p * Jt + t) * nch + c) * 64 + lane) * 8 + e;
                            split_store(v, srow_scale[t * 16 + r], syn_hi, syn_mid, syn_lo, idx);
                        }
                    }
                }
        // analysis: rows = l' per parity, K = rings; values ΔΩ_j λ_l^m(x_j)
        if (JC == 0) continue;
        for (int p = 0; p < 2; ++p)
            for (int lt = 0; lt < nt; ++lt)
                for (int r = 0; r < 16; ++r) {
                    const int lq = 16 * lt + r, l = m + 2 * lq + p;
                    if (l >= lmax) continue;
                    double s = 1.0;
                    if (mode == SHT_PREC_FP16X3) {
                        double amax = 0.0;
                        for (int j = j0; j < n; ++j) amax = std::max(amax, std::fabs(dOmega[j] * LAM(l, j)));
                        s = fp16_scale(amax);
                        ana_sinv[ana_soff[m] + (p * nt + lt) * 16 + r] = (float)(1.0 / s);
                    }
                    for (int jp = 0; jp < J; ++jp) {
                        const int j = j0 + jp;
                        const double v = dOmega[j] * LAM(l, j);
                        if (!split) {
                            ana_f32[ana_off[m] + ((long long)(p * nt + lt) * (JC * sht::kChunk) + jp) * 16 + r] = (float)v;
                        } else {
                            const int c = jp / sht::kChunk, kq = jp % sht::kChunk;
                            const int grp = kq / 8, e = kq % 8, lane = grp * 16 + r;
                            const long long idx = ana_off[m] + (((long long)(p * nt + lt) * JC + c) * 64 + lane) * 8 + e;
                            split_store(v, s, ana_hi, ana_mid, ana_lo, idx);
                        }
                    }
                }
    }

    // ------------------------------------------------------------- FFT tables
    // every factorisation policy's tables (the default one must exist)
    std::vector<int> fac_v[kPolicies], nfac_v[kPolicies];
    const int pol_default = policy_index(FFTPolicy());
    for (int q = 0; q < kPolicies; ++q) {
        rc = fft_tables(g, policy_of(q), fac_v[q], nfac_v[q], pl->h_ring_cost_v[q], pl->h_ring_order_v[q]);
        pl->pol_ok[q] = rc == SHT_OK;
        if (rc && q == pol_default) return rc;
    }
    pl->pol = pol_default;
    std::vector<long long> tw_off(n);
    long long tw_total = 0;
    for (int j = 0; j < n; ++j) {
        tw_off[j] = tw_total;
        tw_total += g.nlon[j];
    }
    std::vector<float> tw(2 * tw_total);
    for (int j = 0; j < n; ++j) {
        const int N = g.nlon[j];
        for (int k = 0; k < N; ++k) {
            const double a = -2.0 * M_PI * (double)k / (double)N;
            tw[2 * (tw_off[j] + k)] = (float)std::cos(a);
            tw[2 * (tw_off[j] + k) + 1] = (float)std::sin(a);
        }
    }
    if (g.nlon_max > sht::kFFTMaxRing) return fail(SHT_ERR_UNSUPPORTED, "ring too long for the LDS FFT");
    const int kg = sht::kFFTCols;
    P.kg = kg;
    P.fft_smem = (g.nlon_max + kg * g.nlon_max) * 8;  // twiddles + one in-place buffer

    // --------------------------------------------------------------- tiles
    std::vector<Tile> syn_t, ana_t;
    std::vector<double> syn_c, ana_c;
    for (int m = 0; m < mmax; ++m) {
        const int J = n - g.j0[m], L = lmax - m;
        for (int b = 0; b * sht::kRingsPerWG < J; ++b) {
            syn_t.push_back({0, m, b, 0});
            syn_c.push_back((double)std::min(sht::kRingsPerWG, J - b * sht::kRingsPerWG) * L);
        }
        for (int b = 0; b * adeg < L; ++b) {
            ana_t.push_back({0, m, b, 0});
            ana_c.push_back((double)std::min(adeg, L - b * adeg) * std::max(J, 1));
        }
    }
    auto sort_tiles = [](std::vector<Tile> &t, std::vector<double> &c) {
        std::vector<size_t> idx(t.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return c[a] > c[b]; });
        std::vector<Tile> out(t.size());
        std::vector<double> oc(t.size());
        for (size_t i = 0; i < idx.size(); ++i) {
            out[i] = t[idx[i]];
            oc[i] = c[idx[i]];
        }
        t.swap(out);
        c.swap(oc);
    };
    sort_tiles(syn_t, syn_c);
    sort_tiles(ana_t, ana_c);
    pl->h_syn = syn_t;
    pl->h_ana = ana_t;
    pl->h_syn_cost = syn_c;
    pl->h_ana_cost = ana_c;

    // ---------------------------------------------------- UV_from_vordiv factors
    std::vector<float> uvx(g.nharm), uv1(g.nharm), uv2(g.nharm);
    std::vector<int> l_of(g.nharm);
    std::vector<float4> gyv(g.nharm);
    std::vector<float2> gxv(g.nharm);
    auto eps = [](double l, double m) { return l > 0 ? std::sqrt((l * l - m * m) / (4.0 * l * l - 1.0)) : 0.0; };
    for (int m = 0; m < mmax; ++m)
        for (int l = m; l < lmax; ++l) {
            const long long i = g.lm_off[m] + (l - m);
            l_of[i] = l;
            uvx[i] = l > 0 ? (float)(-(double)m / ((double)l * (l + 1))) : 0.f;
            uv1[i] = l > 0 ? (float)(eps(l, m) / l) : 0.f;
            uv2[i] = l < lmax - 1 ? (float)(eps(l + 1, m) / (l + 1)) : 0.f;
            // gradient operator factors (gradient_arrays.jl:41-47,55-62,88-95), 0-based l, m
            const bool top = l == lmax - 1;
            gyv[i] = make_float4(top ? 0.f : (float)((l + 1) * eps(l, m)), top ? 0.f : (float)(l * eps(l + 1, m)),
                                 (float)(-(l - 1) * eps(l, m)), top ? 0.f : (float)((l + 2) * eps(l + 1, m)));
            gxv[i] = make_float2(top ? 0.f : (float)m, (float)m);
        }

    // --------------------------------------------------------------- upload
    std::vector<float> coslat_f(coslat_inv.begin(), coslat_inv.end());
    P.n = n; P.lmax = lmax; P.mmax = mmax; P.nharm = g.nharm; P.npoints = g.npoints;
    if ((rc = dev_upload(A, g.nlon, &P.nlon))) return rc;
    if ((rc = dev_upload(A, g.mt, &P.mt))) return rc;
    if ((rc = dev_upload(A, g.mte, &P.mte))) return rc;
    if ((rc = dev_upload(A, g.fboff, &P.fboff))) return rc;
    if ((rc = dev_upload(A, g.ring_n, &P.ring_n))) return rc;
    if ((rc = dev_upload(A, g.ring_s, &P.ring_s))) return rc;
    if ((rc = dev_upload(A, coslat_f, &P.coslat_inv))) return rc;
    std::vector<float> domega_f(dOmega.begin(), dOmega.end());
    if ((rc = dev_upload(A, domega_f, &P.domega))) return rc;
    if ((rc = dev_upload(A, g.j0, &P.j0))) return rc;
    if ((rc = dev_upload(A, g.lm_off, &P.lm_off))) return rc;
    if ((rc = dev_upload(A, syn_off, &P.syn_off))) return rc;
    if ((rc = dev_upload(A, ana_off, &P.ana_off))) return rc;
    if ((rc = dev_upload(A, syn_soff, &P.syn_soff))) return rc;
    if ((rc = dev_upload(A, ana_soff, &P.ana_soff))) return rc;
    if (split) {
        const uint16_t *p16;
        if ((rc = dev_upload(A, syn_hi, &p16))) return rc;
        P.syn_hi = p16;
        if ((rc = dev_upload(A, syn_lo, &p16))) return rc;
        P.syn_lo = p16;
        if ((rc = dev_upload(A, ana_hi, &p16))) return rc;
        P.ana_hi = p16;
        if ((rc = dev_upload(A, ana_lo, &p16))) return rc;
        P.ana_lo = p16;
        if (three) {
            if ((rc = dev_upload(A, syn_mid, &p16))) return rc;
            P.syn_mid = p16;
            if ((rc = dev_upload(A, ana_mid, &p16))) return rc;
            P.ana_mid = p16;
        }
        if ((rc = dev_upload(A, syn_sinv, &P.syn_sinv))) return rc;
        if ((rc = dev_upload(A, ana_sinv, &P.ana_sinv))) return rc;
    } else {
        if ((rc = dev_upload(A, syn_f32, &P.syn_f32))) return rc;
        if ((rc = dev_upload(A, ana_f32, &P.ana_f32))) return rc;
    }
    if ((rc = dev_upload(A, tw, &P.twiddle))) return rc;
    std::vector<long long> poff;  // MFMA prime tables: offset per radix
    {
        std::vector<uint16_t> ptab;
        prime_tables(fac_v, nfac_v, pl->pol_ok, n, ptab, poff);
        if ((rc = dev_upload(A, ptab, &P.prime_tab))) return rc;
        if ((rc = dev_upload(A, poff, &P.prime_off))) return rc;
    }
    if ((rc = dev_upload(A, tw_off, &P.tw_off))) return rc;
    {
        std::vector<float> t32;
        std::vector<uint16_t> t16;
        std::vector<long long> o32, o16;
        dft_tables(g, split, t32, o32, t16, o16);
        if ((rc = dev_upload(A, t32, &P.dft32))) return rc;
        if ((rc = dev_upload(A, o32, &P.dft32_off))) return rc;
        if (split) {
            if ((rc = dev_upload(A, t16, &P.dft16))) return rc;
            if ((rc = dev_upload(A, o16, &P.dft16_off))) return rc;
        }
        for (int syn = 0; syn < 2; ++syn)
            for (int f16 = 0; f16 < 2; ++f16)
                for (int c8 = 0; c8 < 2; ++c8) {
                    int mx = 0;
                    for (int j = 0; j < n; ++j)
                        mx = std::max(mx, sht::dft_smem_bytes(syn != 0, f16 != 0, c8 ? 8 : 4, g.nlon[j] / 4, g.mt[j]));
                    pl->dft_smem[syn][f16][c8] = mx;
                    if (mx > 65536)
                        for (bool b : {false, true})
                            for (bool acc : {false, true})
                                (void)hipFuncSetAttribute(
                                    sht::dft_kernel(syn ? sht::kFFTSynthesis : sht::kFFTAnalysis, f16 != 0, c8 ? 8 : 4, b,
                                                    acc),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, mx);
                }
        bytes_dft = (long long)(t32.size() * 4 + t16.size() * 2);
    }
    if ((rc = dev_upload(A, uvx, &P.uv_x))) return rc;
    if ((rc = dev_upload(A, uv1, &P.uv_1))) return rc;
    if ((rc = dev_upload(A, uv2, &P.uv_2))) return rc;
    if ((rc = dev_upload(A, l_of, &P.l_of))) return rc;
    if ((rc = dev_upload(A, gyv, &P.gy))) return rc;
    if ((rc = dev_upload(A, gxv, &P.gx))) return rc;
    {
        // resident workgroups of each Legendre kernel on the device (for the light-first orders)
        long long slots_syn = 0, slots_ana = 0;
        {
            int cus = 0, bs = 0, ba = 0;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d.device) == hipSuccess &&
                hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &bs, sht::legendre_kernel(sht::kLegendreSynthesis, mode, false), 256, 0) == hipSuccess &&
                hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &ba, sht::legendre_kernel(sht::kLegendreAnalysis, mode, false), 256, 0) == hipSuccess) {
                slots_syn = (long long)cus * bs;
                slots_ana = (long long)cus * ba;
            }
        }
        const int cgs_max = (d.nlayers + sht::kColsComplex - 1) / sht::kColsComplex;
        for (int cgs = 1; cgs <= cgs_max; ++cgs)
            for (int variant = 0; variant < 4; ++variant) {
                const bool grouped = (variant & 1) != 0, light = (variant & 2) != 0;
                std::vector<LaunchUnit> su = legendre_units(syn_t, syn_c, 0, cgs);
                std::vector<LaunchUnit> au = legendre_units(ana_t, ana_c, 0, cgs);
                std::vector<Tile> so, ao;
                if (grouped) {
                    so = xcd_order(su, light ? slots_syn / sht::kXCDs : 0);
                    ao = xcd_order(au, light ? slots_ana / sht::kXCDs : 0);
                } else {
                    // tiles heaviest first, one after another
                    for (auto *uo : {&su, &au}) {
                        std::vector<std::pair<double, Tile>> t;
                        for (const LaunchUnit &u : *uo)
                            for (const Tile &x : u.tiles) t.push_back({u.cost / u.tiles.size(), x});
                        std::stable_sort(t.begin(), t.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
                        std::vector<Tile> &o = uo == &su ? so : ao;
                        for (auto &x : t) o.push_back(x.second);
                    }
                    if (light) {
                        light_first(so, slots_syn);
                        light_first(ao, slots_ana);
                    }
                }
                const Tile *ds = nullptr, *da = nullptr;
                if ((rc = dev_upload(A, so, &ds))) return rc;
                if ((rc = dev_upload(A, ao, &da))) return rc;
                pl->syn_tiles[variant].push_back(ds);
                pl->ana_tiles[variant].push_back(da);
                pl->nsyn[variant].push_back((int)so.size());
                pl->nana[variant].push_back((int)ao.size());
            }
    }
    {
        std::vector<DevPlan> views(kPolicies, P);
        for (int q = 0; q < kPolicies; ++q) {
            if (!pl->pol_ok[q]) continue;
            if ((rc = dev_upload(A, pl->h_ring_order_v[q], &views[q].ring_order))) return rc;
            // the rings' records in launch order
            std::vector<sht::FFTRing> rings(n);
            for (int r = 0; r < n; ++r) {
                const int j = pl->h_ring_order_v[q][r];
                sht::FFTRing &x = rings[r];
                x = sht::FFTRing{};
                x.j = j;
                x.N = g.nlon[j];
                x.mt = g.mt[j];
                x.mte = g.mte[j];
                x.nf = nfac_v[q][j];
                const int R0 = x.nf > 0 ? fac_v[q][(size_t)j * sht::kMaxFactors] : 0;
                // an odd first radix of 11 or more is a prime placed first for the MFMA pass
                // (register passes of 11 and 13 come last; 16 is the only even radix above 8)
                if (R0 >= sht::kMfmaMinPrime && (R0 & 1)) {
                    x.prime = R0;
                    x.M0 = x.N / R0;
                    x.mulM0 = sht::fastdiv_mul(x.M0);
                    x.mulR0 = sht::fastdiv_mul(R0);
                    x.prime_off = poff[R0];
                }
                x.tw_off = tw_off[j];
                x.fboff = g.fboff[j];
                x.ring_n = g.ring_n[j];
                x.ring_s = g.ring_s[j];
                x.passes = (long long)j * sht::kMaxFactors;
            }
            if ((rc = dev_upload(A, rings, &views[q].rings))) return rc;
            if ((rc = dev_upload(A, pass_table(g, fac_v[q], nfac_v[q]), &views[q].passes))) return rc;
            if ((rc = dev_upload(A, nfac_v[q], &views[q].nfac))) return rc;
            pl->Pv[q] = views[q];
        }
        if ((rc = dev_upload(A, views, &pl->dPv))) return rc;
        pl->P = pl->Pv[pl->pol];
        pl->dP = pl->dPv + pl->pol;
    }
    if (P.fft_smem > 65536) {
        for (bool b : {false, true})
            for (int st : {sht::kFFTSynthesis, sht::kFFTAnalysis})
                (void)hipFuncSetAttribute(sht::fft_kernel(st, b), hipFuncAttributeMaxDynamicSharedMemorySize,
                                          P.fft_smem);
        (void)hipFuncSetAttribute(sht::fft_kernel(sht::kFFTSynthesis, false, true),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, P.fft_smem);
    }
    const size_t fb_bytes = (size_t)g.fb_rows * d.nlayers * 16;
    pl->fb_bytes = (fb_bytes + 255) & ~(size_t)255;

    long long bytes = (long long)fb_bytes + bytes_dft;
    bytes += split ? 2LL * sht::split_parts(mode) * (syn_total + ana_total) + 4 * (syn_stotal + ana_stotal)
                   : 4 * (syn_total + ana_total);
    bytes += 8 * tw_total + 16LL * g.nharm + 64LL * n + 48LL * mmax;
    pl->device_bytes = bytes;
    return SHT_OK;
}

}  // namespace

// ================================================================== C ABI

extern "C" {

const char *sht_last_error(void) { return g_err.c_str(); }

const char *sht_version(void) { return "libsht 0.1.0 gfx950"; }

int sht_plan_sizes(const sht_desc *desc, sht_sizes *sizes) {
    if (!sizes) return fail(SHT_ERR_ARG, "null sizes");
    Geometry g;
    int rc = make_geometry(desc, g);
    if (rc) return rc;
    sizes->npoints = g.npoints;
    sizes->nharm = g.nharm;
    sizes->nlat = g.nlat;
    sizes->nlon_max = g.nlon_max;
    sizes->nfreq_total = g.fb_rows;
    // Λ (two directions, padded) + scratch, estimated without building
    long long lam = 0;
    const int adeg = sht::ana_degs_per_wg(desc->precision);
    for (int m = 0; m < g.mmax; ++m) {
        const int J = g.n - g.j0[m], L = g.lmax - m, Le = (L + 1) / 2;
        lam += 2LL * ((J + 15) / 16) * ((Le + 31) / 32) * 512;
        lam += 2LL * ((L + adeg - 1) / adeg) * (adeg / 32) * ((J + 31) / 32) * 512;
    }
    const int lam_bytes = desc->precision == SHT_PREC_FP32 ? 4 : 2 * sht::split_parts(desc->precision);
    sizes->device_bytes = lam * lam_bytes + g.fb_rows * (long long)desc->nlayers * 16;
    return SHT_OK;
}

// Dense transform matrices of a plan (MatrixSpectralTransform, SpeedyTransforms/src/
// matrix_transform.jl:56-172, there probed column by column from the transform itself; here
// the same linear maps written out in FP64): for grid point i of ring j (nlon N, northern
// ring jn = min(j, 2n-1-j), parity sign s = (-1)^(l-m) on southern rings) and every (l, m)
// with m <= mt_jn,
//   B[p, lm] = s λ_lm(x_jn) w_m e^{+2π i m i/N}, w_m = 1 for m = 0 and m = N/2, 2 otherwise
//              (brfft: Im of DC and Nyquist ignored, which B's zero imaginary part there gives)
//   F[lm, p] = s ΔΩ_jn λ_lm(x_jn) e^{-2π i m i/N}            (rfft, then the forward Legendre)
static int build_matrices(sht_plan *pl) {
    const Geometry &g = pl->g;
    const int n = g.n, lmax = g.lmax, mmax = g.mmax;
    const long long np = g.npoints, nh = g.nharm;
    const double bytes = 2.0 * (double)np * 2.0 * (double)nh * 4.0;
    if (bytes > 4.0 * 1024 * 1024 * 1024)
        return fail(SHT_ERR_UNSUPPORTED, "dense transform matrices would take " +
                                             std::to_string((long long)(bytes / 1048576)) + " MiB (limit 4 GiB)");
    std::vector<double> x, w;
    gauss_legendre(2 * n, x, w);
    // both matrices in the GEMM's fragment order (sht::dense_fragment_index)
    const long long S4s = (2 * nh + 15) / 16, S4a = (np + 15) / 16;
    std::vector<float> B((size_t)((np + 15) / 16) * 16 * S4s * 16, 0.f), F((size_t)((2 * nh + 15) / 16) * 16 * S4a * 16, 0.f),
        ci(np);
    std::vector<double> lam(nh);
    for (int jn = 0; jn < n; ++jn) {
        const double cosc = -x[jn], sinc = std::sqrt(std::max(0.0, 1.0 - x[jn] * x[jn]));
        const double dO = w[jn] * 2.0 * M_PI / g.nlon[jn];
        if (sht_legendre_column(lmax, mmax, cosc, lam.data()) != SHT_OK) return SHT_ERR_ARG;
        const int N = g.nlon[jn];
        for (int hemi = 0; hemi < 2; ++hemi) {
            const long long p0 = hemi == 0 ? g.ring_n[jn] : g.ring_s[jn];
            for (int i = 0; i < N; ++i) ci[p0 + i] = (float)(1.0 / sinc);
            for (int m = 0; m <= g.mt[jn]; ++m) {
                const double wm = (m == 0 || 2 * m == N) ? 1.0 : 2.0;
                for (int i = 0; i < N; ++i) {
                    const double ph = 2.0 * M_PI * (double)((long long)m * i % N) / N;
                    const double c = std::cos(ph), sn = std::sin(ph);
                    const long long p = p0 + i;
                    for (int l = m; l < lmax; ++l) {
                        const long long lm = g.lm_off[m] + (l - m);
                        const double v = (hemi == 1 && ((l - m) & 1)) ? -lam[lm] : lam[lm];
                        B[sht::dense_fragment_index(p, 2 * lm, S4s)] = (float)(v * wm * c);
                        B[sht::dense_fragment_index(p, 2 * lm + 1, S4s)] = (float)(-v * wm * sn);
                        F[sht::dense_fragment_index(2 * lm, p, S4a)] = (float)(dO * v * c);
                        F[sht::dense_fragment_index(2 * lm + 1, p, S4a)] = (float)(-dO * v * sn);
                    }
                }
            }
        }
    }
    const float *d = nullptr;
    int rc;
    if ((rc = dev_upload(pl->allocs, B, &d))) return rc;
    pl->mat_syn = const_cast<float *>(d);
    if ((rc = dev_upload(pl->allocs, F, &d))) return rc;
    pl->mat_ana = const_cast<float *>(d);
    if ((rc = dev_upload(pl->allocs, ci, &d))) return rc;
    pl->pt_coslat_inv = const_cast<float *>(d);
    // workspace of the split-K GEMMs: the largest over both directions and every K
    long long work = 1;
    for (int K = 1; K <= pl->d.nlayers; ++K) {
        const sht::DenseSplit ds = sht::dense_split(np, 2 * nh, K), da = sht::dense_split(2 * nh, np, K);
        if (ds.s > 1) work = std::max(work, (long long)ds.s * K * np);
        if (da.s > 1) work = std::max(work, (long long)da.s * K * 2 * nh);
    }
    pl->dense_bytes = ((size_t)work * 4 + 255) & ~(size_t)255;  // in each stream's scratch
    pl->device_bytes += (long long)bytes + np * 4 + work * 4;
    pl->matrix = true;
    return SHT_OK;
}

// A regular plan, or with matrix the dense one, with the null stream's scratch allocated.
static int create_plan(const sht_desc *desc, bool matrix, sht_plan **plan) {
    sht::TraceRange trace_range(matrix ? "sht_plan_create_matrix" : "sht_plan_create");
    if (!plan) return fail(SHT_ERR_ARG, "null plan pointer");
    *plan = nullptr;
    sht_plan *pl = new (std::nothrow) sht_plan();
    if (!pl) return fail(SHT_ERR_ALLOC, "host allocation failed");
    int rc = make_geometry(desc, pl->g);
    if (rc) { delete pl; return rc; }
    pl->d = *desc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        delete pl;
        return fail(SHT_ERR_HIP, "no HIP device available");
    }
    if (desc->device < 0 || desc->device >= ndev) {
        delete pl;
        return fail(SHT_ERR_ARG, "device ordinal out of range");
    }
    DeviceGuard guard(desc->device);
    rc = build_plan(pl);
    if (!rc && matrix) rc = build_matrices(pl);
    if (!rc && !stream_scratch(pl, nullptr)) rc = SHT_ERR_ALLOC;
    if (!rc) rc = sht::static_scratch_setup(desc->device);
    if (rc) { free_plan(pl); return rc; }
    *plan = pl;
    return SHT_OK;
}

int sht_plan_create(const sht_desc *desc, sht_plan **plan) { return create_plan(desc, false, plan); }

int sht_plan_create_matrix(const sht_desc *desc, sht_plan **plan) { return create_plan(desc, true, plan); }

int sht_plan_destroy(sht_plan *plan) {
    sht::TraceRange trace_range("sht_plan_destroy");
    if (!plan) return SHT_OK;
    DeviceGuard guard(plan->d.device);
    (void)hipDeviceSynchronize();
    free_plan(plan);
    return SHT_OK;
}

// Blocks of a single-call FFT launch (the kernels map block -> ring, column group):
// paired, all column groups of a ring on one XCD (ring ranks padded to a multiple of the
// XCD count); otherwise one block per (ring, column group) in rank order.
static unsigned fft_blocks(int nrings, int K, bool paired) {
    const int kgs = (K + sht::kFFTCols - 1) / sht::kFFTCols;
    if (!paired) return (unsigned)(nrings * kgs);
    return (unsigned)(sht::kXCDs * kgs * ((nrings + sht::kXCDs - 1) / sht::kXCDs));
}

// Launches one stage kernel (every stage kernel has the same signature).
static hipError_t launch_stage(const void *func, dim3 grid, unsigned smem, const DevPlan *plans, const CallArgs &c,
                               const CallArgs *calls, const Tile *tiles, void *stream) {
    const DevPlan *a0 = plans;
    CallArgs a1 = c;
    const CallArgs *a2 = calls;
    const Tile *a3 = tiles;
    void *args[] = {&a0, &a1, &a2, &a3};
    return hipLaunchKernel(func, grid, dim3(256), args, smem, reinterpret_cast<hipStream_t>(stream));
}

static int check_call(sht_plan *plan, const void *a, const void *b, int K) {
    if (!plan) return fail(SHT_ERR_ARG, "null plan");
    if (!a || !b) return fail(SHT_ERR_ARG, "null data pointer");
    if (K < 1 || K > plan->d.nlayers)
        return fail(SHT_ERR_DIM, "DimensionMismatch: K=" + std::to_string(K) + " columns, plan holds nlayers=" +
                                     std::to_string(plan->d.nlayers));
    return SHT_OK;
}

// A fused spectral Laplacian of a transform call (CallArgs::lap and its factors).
struct LapOp {
    int lap = 0;  // 0 none, 1 ∇², 2 ∇⁻²
    float sign = 1.f, scale = 1.f;
};

// Launches one stage of a single call (c: its arguments, K columns) on the plan's chosen
// launch orders, bracketed by profiling events when profiling is on.
// Kernel, grid, dynamic LDS and tile list of one stage of a single call on the plan's chosen
// launch orders.
struct StageLaunch {
    const void *func;
    dim3 grid;
    unsigned smem;
    const Tile *tiles;
};

static StageLaunch plan_stage_launch(const sht_plan *plan, int stage, const CallArgs &c) {
    const int K = c.K;
    const int cgs = (K + sht::kColsComplex - 1) / sht::kColsComplex;
    const int gx = plan->leg_xcd;
    switch (stage) {
        case sht::kLegendreSynthesis:
            return {sht::legendre_kernel(stage, plan->d.precision, false), dim3(plan->nsyn[gx][cgs - 1]), 0,
                    plan->syn_tiles[gx][cgs - 1]};
        case sht::kFFTSynthesis:
        case sht::kFFTAnalysis: {
            if (plan->fourier != sht::kFourierRadix) {  // ring DFT on MFMA: blocks (ring rank, column group)
                const bool f16 = plan->fourier == sht::kFourierDft16;
                const int cg = plan->dft_cg;
                const bool syn = stage == sht::kFFTSynthesis;
                return {sht::dft_kernel(stage, f16, cg, false, syn && c.accumulate != 0),
                        dim3(plan->g.n * ((K + cg - 1) / cg)), (unsigned)plan->dft_smem[syn][f16][cg == 8], nullptr};
            }
            if (stage == sht::kFFTSynthesis)
                return {sht::fft_kernel(stage, false, c.accumulate != 0), dim3(fft_blocks(plan->g.n, K, c.paired != 0)),
                        (unsigned)plan->P.fft_smem, reinterpret_cast<const Tile *>(plan->P.rings)};
            // the FFT kernels of a single call take the ring records in place of a tile list
            return {sht::fft_kernel(stage, false), dim3(fft_blocks(plan->g.n, K, false)), (unsigned)plan->P.fft_smem,
                    reinterpret_cast<const Tile *>(plan->P.rings)};
        }
        default:
            return {sht::legendre_kernel(stage, plan->d.precision, false), dim3(plan->nana[gx][cgs - 1]), 0,
                    plan->ana_tiles[gx][cgs - 1]};
    }
}

// Launches one stage of a single call, bracketed by profiling events when profiling is on.
static hipError_t launch_plan_stage(sht_plan *plan, int stage, const CallArgs &c, void *stream) {
    static const int kernel_id[4] = {SHT_KERNEL_LEGENDRE_SYNTHESIS, SHT_KERNEL_FFT_SYNTHESIS, SHT_KERNEL_FFT_ANALYSIS,
                                     SHT_KERNEL_LEGENDRE_ANALYSIS};
    ProfScope ps(plan, kernel_id[stage], stream);
    const StageLaunch l = plan_stage_launch(plan, stage, c);
    return launch_stage(l.func, l.grid, l.smem, plan->dP, c, nullptr, l.tiles, stream);
}

constexpr int kMaxGraphs = 64;  // cached graphs per direction and plan

// Runs the two stages of a call (s0 then s1) from the plan's graph cache when graphs are on:
// replays the call's graph, or builds and instantiates it (explicit kernel nodes, no stream
// capture) while fewer than kMaxGraphs are cached.  Returns false when the call is to be
// launched directly (graphs off, profiling on, a capturing stream, a full cache, or a
// failed graph); e = the launch status otherwise.
static bool launch_from_graph(sht_plan *plan, int dir, int s0, int s1, const CallArgs &c, void *stream,
                              hipError_t &e) {
    if (!plan->graphs || plan->prof.on) return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(reinterpret_cast<hipStream_t>(stream), &cs) != hipSuccess ||
        cs != hipStreamCaptureStatusNone)
        return false;
    // key: the call's fields one by one (CallArgs has tail padding) and the tuning
    std::string key;
    auto put = [&key](const auto &v) { key.append(reinterpret_cast<const char *>(&v), sizeof(v)); };
    put(c.K), put(c.unscale), put(c.paired), put(c.in), put(c.out), put(c.ld_in), put(c.ld_out), put(c.fbuf);
    put(c.adj), put(c.accumulate), put(c.lap), put(c.lap_sign), put(c.lap_scale);
    put(plan->pol), put(plan->leg_xcd), put(plan->fft_pair), put(plan->fourier), put(plan->dft_cg);
    hipGraphExec_t exec = nullptr;
    {
        std::lock_guard<std::mutex> lock(plan->graph_mutex);
        auto &cache = plan->graph_cache[dir];
        auto it = cache.find(key);
        if (it != cache.end()) {
            if (!it->second) return false;  // its graph failed: direct launches
            exec = it->second;
        } else {
            if ((int)cache.size() >= kMaxGraphs) return false;
            hipGraph_t graph = nullptr;
            bool ok = hipGraphCreate(&graph, 0) == hipSuccess;
            hipGraphNode_t prev = nullptr;
            for (int st : {s0, s1}) {
                if (!ok) break;
                const StageLaunch l = plan_stage_launch(plan, st, c);
                const DevPlan *a0 = plan->dP;
                CallArgs a1 = c;
                const CallArgs *a2 = nullptr;
                const Tile *a3 = l.tiles;
                void *args[] = {&a0, &a1, &a2, &a3};
                hipKernelNodeParams kp{};
                kp.func = const_cast<void *>(l.func);
                kp.gridDim = l.grid;
                kp.blockDim = dim3(256);
                kp.sharedMemBytes = l.smem;
                kp.kernelParams = args;
                kp.extra = nullptr;
                hipGraphNode_t node;
                ok = hipGraphAddKernelNode(&node, graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp) == hipSuccess;
                prev = node;
            }
            if (ok && !plan->graph_force_failure)
                ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
            else
                ok = false;
            if (graph) (void)hipGraphDestroy(graph);
            if (!ok) {
                (void)hipGetLastError();
                exec = nullptr;
                ++plan->graph_failed;
            }
            cache[key] = exec;
            if (!exec) return false;
        }
    }
    e = hipGraphLaunch(exec, reinterpret_cast<hipStream_t>(stream));
    return true;
}

// Scratch of a call: the caller's (sized for the plan), else the stream's own.
static void *call_scratch(sht_plan *plan, void *scratch, void *stream) {
    return scratch ? scratch : stream_scratch(plan, stream);
}

// Spectral -> grid stages (Legendre synthesis, FFT synthesis); adj = 2 and accumulate turn
// them into the pullback of the analysis; op fuses a Laplacian of the input; scratch (if
// not null) replaces the stream's scratch.
static int run_synthesis(sht_plan *plan, const void *coeffs, int64_t ld_spec, void *field, int64_t ld_grid, int32_t K,
                         int32_t unscale_coslat, int adj, int accumulate, void *stream, LapOp op = LapOp(),
                         void *scratch = nullptr) {
    int rc = check_call(plan, coeffs, field, K);
    if (rc) return rc;
    if (ld_spec < plan->g.nharm) return fail(SHT_ERR_DIM, "DimensionMismatch: ld_spec < nharm");
    if (ld_grid < plan->g.npoints) return fail(SHT_ERR_DIM, "DimensionMismatch: ld_grid < npoints");
    DeviceGuard guard(plan->d.device);
    void *ws = call_scratch(plan, scratch, stream);
    if (!ws) return SHT_ERR_ALLOC;
    if (plan->matrix && adj == 0 && !accumulate && op.lap == 0) {  // field = B_int a_int, one GEMM (rows / cos(lat))
        float *work = static_cast<float *>(ws);  // the GEMM needs no Fourier buffer
        rc = sht::launch_dense_gemm(plan->mat_syn, plan->g.npoints, 2 * plan->g.nharm,
                                    static_cast<const float *>(coeffs), 2 * ld_spec, K, static_cast<float *>(field),
                                    ld_grid, unscale_coslat ? plan->pt_coslat_inv : nullptr, work, stream);
        if (rc) return fail(SHT_ERR_HIP, std::string("dense synthesis launch: ") + hipGetErrorString((hipError_t)rc));
        return SHT_OK;
    }
    const CallArgs c{0, K, unscale_coslat ? 1 : 0, plan->fft_pair, coeffs, field, ld_spec, ld_grid, ws,
                     adj, accumulate ? 1 : 0, op.lap, op.sign, op.scale};
    hipError_t e;
    if (launch_from_graph(plan, 0, sht::kLegendreSynthesis, sht::kFFTSynthesis, c, stream, e)) {
        if (e != hipSuccess) return fail(SHT_ERR_HIP, std::string("synthesis graph launch: ") + hipGetErrorString(e));
        return SHT_OK;
    }
    e = launch_plan_stage(plan, sht::kLegendreSynthesis, c, stream);
    if (e != hipSuccess) return fail(SHT_ERR_HIP, std::string("legendre synthesis launch: ") + hipGetErrorString(e));
    e = launch_plan_stage(plan, sht::kFFTSynthesis, c, stream);
    if (e != hipSuccess) return fail(SHT_ERR_HIP, std::string("fft synthesis launch: ") + hipGetErrorString(e));
    return SHT_OK;
}

// Grid -> spectral stages (FFT analysis, Legendre analysis); adj = 1 and accumulate turn
// them into the pullback of the synthesis; op fuses a Laplacian of the output.
static int run_analysis(sht_plan *plan, const void *field, int64_t ld_grid, void *coeffs, int64_t ld_spec, int32_t K,
                        int32_t unscale_coslat, int adj, int accumulate, void *stream, LapOp op = LapOp(),
                        void *scratch = nullptr) {
    int rc = check_call(plan, field, coeffs, K);
    if (rc) return rc;
    if (ld_spec < plan->g.nharm) return fail(SHT_ERR_DIM, "DimensionMismatch: ld_spec < nharm");
    if (ld_grid < plan->g.npoints) return fail(SHT_ERR_DIM, "DimensionMismatch: ld_grid < npoints");
    DeviceGuard guard(plan->d.device);
    void *ws = call_scratch(plan, scratch, stream);
    if (!ws) return SHT_ERR_ALLOC;
    if (plan->matrix && adj == 0 && !accumulate && op.lap == 0) {  // a_int = F_int f, one GEMM
        float *work = static_cast<float *>(ws);  // the GEMM needs no Fourier buffer
        rc = sht::launch_dense_gemm(plan->mat_ana, 2 * plan->g.nharm, plan->g.npoints,
                                    static_cast<const float *>(field), ld_grid, K, static_cast<float *>(coeffs),
                                    2 * ld_spec, nullptr, work, stream);
        if (rc) return fail(SHT_ERR_HIP, std::string("dense analysis launch: ") + hipGetErrorString((hipError_t)rc));
        return SHT_OK;
    }
    const CallArgs c{0,  K,   unscale_coslat ? 1 : 0, 0,      field,   coeffs, ld_grid, ld_spec,
                     ws, adj, accumulate ? 1 : 0,     op.lap, op.sign, op.scale};
    hipError_t e;
    if (launch_from_graph(plan, 1, sht::kFFTAnalysis, sht::kLegendreAnalysis, c, stream, e)) {
        if (e != hipSuccess) return fail(SHT_ERR_HIP, std::string("analysis graph launch: ") + hipGetErrorString(e));
        return SHT_OK;
    }
    e = launch_plan_stage(plan, sht::kFFTAnalysis, c, stream);
    if (e != hipSuccess) return fail(SHT_ERR_HIP, std::string("fft analysis launch: ") + hipGetErrorString(e));
    e = launch_plan_stage(plan, sht::kLegendreAnalysis, c, stream);
    if (e != hipSuccess) return fail(SHT_ERR_HIP, std::string("legendre analysis launch: ") + hipGetErrorString(e));
    return SHT_OK;
}

int sht_synthesis(sht_plan *plan, const void *coeffs, int64_t ld_spec, void *field, int64_t ld_grid, int32_t K,
                  int32_t unscale_coslat, void *stream) {
    sht::TraceRange trace_range("sht_synthesis");
    return run_synthesis(plan, coeffs, ld_spec, field, ld_grid, K, unscale_coslat, 0, 0, stream);
}

int sht_analysis(sht_plan *plan, const void *field, int64_t ld_grid, void *coeffs, int64_t ld_spec, int32_t K,
                 void *stream) {
    sht::TraceRange trace_range("sht_analysis");
    return run_analysis(plan, field, ld_grid, coeffs, ld_spec, K, 0, 0, 0, stream);
}

int sht_spec2grid_pullback(sht_plan *plan, const void *field_bar, int64_t ld_grid, void *coeffs_bar, int64_t ld_spec,
                           int32_t K, int32_t unscale_coslat, int32_t accumulate, void *stream) {
    sht::TraceRange trace_range("sht_spec2grid_pullback");
    return run_analysis(plan, field_bar, ld_grid, coeffs_bar, ld_spec, K, unscale_coslat, 1, accumulate, stream);
}

int sht_grid2spec_pullback(sht_plan *plan, const void *coeffs_bar, int64_t ld_spec, void *field_bar, int64_t ld_grid,
                           int32_t K, int32_t accumulate, void *stream) {
    sht::TraceRange trace_range("sht_grid2spec_pullback");
    return run_synthesis(plan, coeffs_bar, ld_spec, field_bar, ld_grid, K, 0, 2, accumulate, stream);
}

}  // extern "C"

size_t sht::plan_scratch_bytes(const sht_plan *plan, int K) {
    return plan->matrix ? plan->dense_bytes : (size_t)plan->g.fb_rows * (size_t)K * 16;
}

int sht::plan_device(const sht_plan *plan) { return plan->d.device; }

int sht::transform_with_scratch(sht_plan *plan, int direction, const void *in, int64_t ld_in, void *out,
                                int64_t ld_out, int32_t K, int32_t unscale_coslat, void *scratch, void *stream) {
    if (direction == SHT_DIR_SYNTHESIS)
        return run_synthesis(plan, in, ld_in, out, ld_out, K, unscale_coslat, 0, 0, stream, LapOp(), scratch);
    if (direction == SHT_DIR_ANALYSIS) {
        if (unscale_coslat) return fail(SHT_ERR_ARG, "unscale_coslat applies to spectral->grid transforms only");
        return run_analysis(plan, in, ld_in, out, ld_out, K, 0, 0, 0, stream, LapOp(), scratch);
    }
    return fail(SHT_ERR_ARG, "direction must be SHT_DIR_SYNTHESIS or SHT_DIR_ANALYSIS");
}

extern "C" {

// The Laplacian's factors on a sphere of the given radius (as sht_laplace: 1/R² for ∇², R²
// for ∇⁻²), or an error status.
static int lap_op(int32_t inverse, int32_t flipsign, double radius, LapOp &op) {
    if (!(radius > 0)) return fail(SHT_ERR_ARG, "radius must be positive");
    const double r2 = radius * radius;
    op.lap = inverse ? 2 : 1;
    op.sign = flipsign ? -1.f : 1.f;
    op.scale = (float)(inverse ? r2 : 1.0 / r2);
    return SHT_OK;
}

int sht_synthesis_laplace(sht_plan *plan, const void *coeffs, int64_t ld_spec, void *field, int64_t ld_grid, int32_t K,
                          int32_t inverse, int32_t flipsign, double radius, int32_t unscale_coslat, void *stream) {
    sht::TraceRange trace_range("sht_synthesis_laplace");
    LapOp op;
    const int rc = lap_op(inverse, flipsign, radius, op);
    if (rc) return rc;
    return run_synthesis(plan, coeffs, ld_spec, field, ld_grid, K, unscale_coslat, 0, 0, stream, op);
}

int sht_analysis_laplace(sht_plan *plan, const void *field, int64_t ld_grid, void *coeffs, int64_t ld_spec, int32_t K,
                         int32_t inverse, int32_t flipsign, int32_t add, double radius, void *stream) {
    sht::TraceRange trace_range("sht_analysis_laplace");
    LapOp op;
    const int rc = lap_op(inverse, flipsign, radius, op);
    if (rc) return rc;
    return run_analysis(plan, field, ld_grid, coeffs, ld_spec, K, 0, 0, add ? 1 : 0, stream, op);
}

int sht_uv_from_vordiv(sht_plan *plan, const void *vor, const void *div, void *U, void *V, int64_t ld, int32_t K,
                       double radius, void *stream) {
    sht::TraceRange trace_range("sht_uv_from_vordiv");
    int rc = check_call(plan, vor, div, K);
    if (rc) return rc;
    if (!U || !V) return fail(SHT_ERR_ARG, "null data pointer");
    if (ld < plan->g.nharm) return fail(SHT_ERR_DIM, "DimensionMismatch: ld < nharm");
    DeviceGuard guard(plan->d.device);
    rc = sht::launch_uv_from_vordiv(plan->dP, plan->g.nharm, vor, div, U, V, ld, K, (float)radius, stream);
    if (rc) return fail(SHT_ERR_HIP, std::string("UV_from_vordiv launch: ") + hipGetErrorString((hipError_t)rc));
    return SHT_OK;
}

int sht_divergence(sht_plan *plan, const void *u, const void *v, void *out, int64_t ld, int32_t K, int32_t curl,
                   int32_t flipsign, int32_t add, double radius, void *stream) {
    sht::TraceRange trace_range("sht_divergence");
    int rc = check_call(plan, u, v, K);
    if (rc) return rc;
    if (!out) return fail(SHT_ERR_ARG, "null data pointer");
    if (out == u || out == v) return fail(SHT_ERR_ARG, "output must not alias an input");
    if (ld < plan->g.nharm) return fail(SHT_ERR_DIM, "DimensionMismatch: ld < nharm");
    if (!(radius > 0)) return fail(SHT_ERR_ARG, "radius must be positive");
    DeviceGuard guard(plan->d.device);
    const int flags = (flipsign ? sht::kOpFlip : 0) | (add ? sht::kOpAdd : 0);
    rc = sht::launch_spectral_op(plan->dP, plan->g.nharm, curl ? sht::kOpCurl : sht::kOpDivergence, u, v, out, nullptr,
                                 ld, K, flags, (float)(1.0 / radius), stream);
    if (rc) return fail(SHT_ERR_HIP, std::string("divergence/curl launch: ") + hipGetErrorString((hipError_t)rc));
    return SHT_OK;
}

int sht_laplace(sht_plan *plan, const void *alms, void *out, int64_t ld, int32_t K, int32_t inverse, int32_t flipsign,
                int32_t add, double radius, void *stream) {
    sht::TraceRange trace_range("sht_laplace");
    int rc = check_call(plan, alms, out, K);
    if (rc) return rc;
    if (ld < plan->g.nharm) return fail(SHT_ERR_DIM, "DimensionMismatch: ld < nharm");
    if (!(radius > 0)) return fail(SHT_ERR_ARG, "radius must be positive");
    DeviceGuard guard(plan->d.device);
    const int flags = (flipsign ? sht::kOpFlip : 0) | (add ? sht::kOpAdd : 0) | (inverse ? sht::kOpInverse : 0);
    const double r2 = radius * radius;
    rc = sht::launch_spectral_op(plan->dP, plan->g.nharm, sht::kOpLaplace, alms, nullptr, out, nullptr, ld, K, flags,
                                 (float)(inverse ? r2 : 1.0 / r2), stream);
    if (rc) return fail(SHT_ERR_HIP, std::string("laplace launch: ") + hipGetErrorString((hipError_t)rc));
    return SHT_OK;
}

int sht_gradient(sht_plan *plan, const void *p, void *dpdx, void *dpdy, int64_t ld, int32_t K, double radius,
                 void *stream) {
    sht::TraceRange trace_range("sht_gradient");
    int rc = check_call(plan, p, dpdx, K);
    if (rc) return rc;
    if (!dpdy) return fail(SHT_ERR_ARG, "null data pointer");
    if (dpdx == p || dpdy == p || dpdx == dpdy) return fail(SHT_ERR_ARG, "outputs must not alias the input or each other");
    if (ld < plan->g.nharm) return fail(SHT_ERR_DIM, "DimensionMismatch: ld < nharm");
    if (!(radius > 0)) return fail(SHT_ERR_ARG, "radius must be positive");
    DeviceGuard guard(plan->d.device);
    rc = sht::launch_spectral_op(plan->dP, plan->g.nharm, sht::kOpGradient, p, nullptr, dpdx, dpdy, ld, K, 0,
                                 (float)(1.0 / radius), stream);
    if (rc) return fail(SHT_ERR_HIP, std::string("gradient launch: ") + hipGetErrorString((hipError_t)rc));
    return SHT_OK;
}

int sht_time_stage(sht_plan *plan, int32_t stage, void *coeffs, int64_t ld_spec, void *field, int64_t ld_grid, int32_t K,
                   int32_t reps, void *stream, double *avg_ms) {
    if (!avg_ms || reps < 1 || stage < 0 || stage > 3) return fail(SHT_ERR_ARG, "bad arguments");
    int rc = check_call(plan, coeffs, field, K);
    if (rc) return rc;
    if (plan->matrix) return fail(SHT_ERR_UNSUPPORTED, "a dense plan has no transform stages");
    if (ld_spec < plan->g.nharm || ld_grid < plan->g.npoints) return fail(SHT_ERR_DIM, "DimensionMismatch: stride");
    DeviceGuard guard(plan->d.device);
    void *ws = stream_scratch(plan, stream);
    if (!ws) return SHT_ERR_ALLOC;
    const bool spec_in = stage == sht::kLegendreSynthesis || stage == sht::kFFTSynthesis;
    const CallArgs c{0, K, 0, spec_in ? plan->fft_pair : 0, spec_in ? coeffs : field, spec_in ? field : coeffs,
                     spec_in ? ld_spec : ld_grid, spec_in ? ld_grid : ld_spec, ws};
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return fail(SHT_ERR_HIP, "hipEventCreate failed");
    }
    hipError_t e = launch_plan_stage(plan, stage, c, stream);  // untimed: first-launch costs
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int i = 0; i < reps && e == hipSuccess; ++i) e = launch_plan_stage(plan, stage, c, stream);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(SHT_ERR_HIP, std::string("stage timing: ") + hipGetErrorString(e));
    *avg_ms = (double)ms / reps;
    return SHT_OK;
}

int sht_profile_begin(sht_plan *plan, int32_t max_calls) {
    if (!plan || max_calls < 1) return fail(SHT_ERR_ARG, "bad arguments");
    DeviceGuard guard(plan->d.device);
    Profile &pr = plan->prof;
    for (hipEvent_t e : pr.ev) (void)hipEventDestroy(e);
    pr = Profile();
    pr.ev.resize((size_t)SHT_NKERNELS * max_calls * 2);
    for (auto &e : pr.ev) {
        if (hipEventCreate(&e) != hipSuccess) return fail(SHT_ERR_HIP, "hipEventCreate failed");
    }
    pr.max_calls = max_calls;
    pr.on = true;
    return SHT_OK;
}

int sht_profile_end(sht_plan *plan, double *avg_ms, int32_t *counts) {
    if (!plan || !avg_ms || !counts) return fail(SHT_ERR_ARG, "bad arguments");
    DeviceGuard guard(plan->d.device);
    Profile &pr = plan->prof;
    int rc = SHT_OK;
    for (int k = 0; k < SHT_NKERNELS; ++k) {
        double tot = 0.0;
        for (int c = 0; c < pr.ncalls[k]; ++c) {
            hipEvent_t a = pr.ev[((size_t)k * pr.max_calls + c) * 2], b = pr.ev[((size_t)k * pr.max_calls + c) * 2 + 1];
            float ms = 0.f;
            if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) {
                rc = fail(SHT_ERR_HIP, "event timing failed");
                break;
            }
            tot += ms;
        }
        counts[k] = pr.ncalls[k];
        avg_ms[k] = pr.ncalls[k] ? tot / pr.ncalls[k] : 0.0;
    }
    for (hipEvent_t e : pr.ev) (void)hipEventDestroy(e);
    pr = Profile();
    return rc;
}

namespace {
// Chooses the FFT tables of a factorisation policy (all are on the device already: only
// the plan's host state changes, launches already queued keep the tables they were given).
int apply_fft_policy(sht_plan *plan, const FFTPolicy &pol) {
    const int q = policy_index(pol);
    if (!plan->pol_ok[q])
        return fail(SHT_ERR_UNSUPPORTED, "the FFT factorisation policy has no kernel for some ring length");
    plan->pol = q;
    plan->P = plan->Pv[q];
    plan->dP = plan->dPv + q;
    return SHT_OK;
}

struct TuneKey {
    int grid, nlat_half, lmax, mmax, nlayers, precision, K;
    bool operator<(const TuneKey &o) const {
        return std::tie(grid, nlat_half, lmax, mmax, nlayers, precision, K) <
               std::tie(o.grid, o.nlat_half, o.lmax, o.mmax, o.nlayers, o.precision, o.K);
    }
};
struct TuneChoice {
    FFTPolicy pol;
    int leg_xcd, fft_pair;
    int fourier, dft_cg;
};
std::map<TuneKey, TuneChoice> g_tune_cache;
std::mutex g_tune_mutex;

// The autotuner also times the ring DFT on MFMA.
constexpr bool kAutotuneDft = false;

// Applies a tuning choice (host state only).
int apply_choice(sht_plan *plan, const TuneChoice &c) {
    if (c.fourier == sht::kFourierDft16 && plan->d.precision == SHT_PREC_FP32)
        return fail(SHT_ERR_UNSUPPORTED, "the FP16 ring DFT needs a split-precision plan");
    const int rc = apply_fft_policy(plan, c.pol);
    if (rc) return rc;
    plan->leg_xcd = c.leg_xcd;
    plan->fft_pair = c.fft_pair;
    plan->fourier = c.fourier;
    plan->dft_cg = c.dft_cg;
    return SHT_OK;
}

// Tuning cache file: one flat JSON object of integers per line, keys as below.
const char *const kTuneFields[] = {"grid", "nlat_half", "lmax", "mmax", "nlayers", "precision", "K",
                                   "mfma_min", "eights", "odd_first", "leg_xcd", "fft_pair", "fourier", "dft_cg"};
constexpr int kTuneNFields = 14;

void tune_to_ints(const TuneKey &k, const TuneChoice &c, long long *v) {
    const long long a[kTuneNFields] = {k.grid, k.nlat_half, k.lmax, k.mmax, k.nlayers, k.precision, k.K,
                                       c.pol.mfma_min, c.pol.eights, c.pol.odd_first, c.leg_xcd, c.fft_pair,
                                       c.fourier, c.dft_cg};
    std::copy(a, a + kTuneNFields, v);
}

// Parses one line; false if a field is missing or a value out of range.
bool tune_from_line(const std::string &line, TuneKey &k, TuneChoice &c) {
    long long v[kTuneNFields];
    for (int f = 0; f < kTuneNFields; ++f) {
        const std::string key = std::string("\"") + kTuneFields[f] + "\"";
        const size_t at = line.find(key);
        if (at == std::string::npos) return false;
        size_t q = line.find(':', at + key.size());
        if (q == std::string::npos) return false;
        char *end = nullptr;
        v[f] = std::strtoll(line.c_str() + q + 1, &end, 10);
        if (end == line.c_str() + q + 1) return false;
    }
    k = TuneKey{(int)v[0], (int)v[1], (int)v[2], (int)v[3], (int)v[4], (int)v[5], (int)v[6]};
    c.pol.mfma_min = (int)v[7];
    c.pol.eights = (int)v[8];
    c.pol.odd_first = v[9] != 0;
    c.leg_xcd = (int)v[10];
    c.fft_pair = (int)v[11];
    c.fourier = (int)v[12];
    c.dft_cg = (int)v[13];
    return (c.pol.mfma_min == 11 || c.pol.mfma_min == 17) && c.pol.eights >= 0 && c.pol.eights <= 2 &&
           c.leg_xcd >= 0 && c.leg_xcd <= 3 && c.fft_pair >= 0 &&
           c.fft_pair <= 1 && c.fourier >= 0 && c.fourier <= 2 && (c.dft_cg == 4 || c.dft_cg == 8) && k.K >= 1;
}
}  // namespace

int sht_tune_cache_save(const char *path) {
    if (!path) return fail(SHT_ERR_ARG, "null path");
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(SHT_ERR_ARG, std::string("cannot write ") + path);
    for (const auto &kv : g_tune_cache) {
        long long v[kTuneNFields];
        tune_to_ints(kv.first, kv.second, v);
        std::fputc('{', f);
        for (int i = 0; i < kTuneNFields; ++i) std::fprintf(f, "%s\"%s\": %lld", i ? ", " : "", kTuneFields[i], v[i]);
        std::fputs("}\n", f);
    }
    const bool ok = std::fclose(f) == 0;
    return ok ? SHT_OK : fail(SHT_ERR_ARG, std::string("cannot write ") + path);
}

int sht_tune_cache_load(const char *path, int32_t *nloaded) {
    if (!path) return fail(SHT_ERR_ARG, "null path");
    FILE *f = std::fopen(path, "r");
    if (!f) return fail(SHT_ERR_ARG, std::string("cannot read ") + path);
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    int n = 0, bad = 0;
    char buf[1024];
    while (std::fgets(buf, sizeof buf, f)) {
        const std::string line(buf);
        if (line.find('{') == std::string::npos) continue;
        TuneKey k;
        TuneChoice c;
        if (tune_from_line(line, k, c)) {
            g_tune_cache[k] = c;
            ++n;
        } else {
            ++bad;
        }
    }
    std::fclose(f);
    if (nloaded) *nloaded = n;
    return bad ? fail(SHT_ERR_ARG, std::to_string(bad) + " malformed line(s) in " + path) : SHT_OK;
}

int sht_plan_autotune(sht_plan *plan, int32_t K, void *stream) {
    sht::TraceRange trace_range("sht_plan_autotune");
    if (!plan) return fail(SHT_ERR_ARG, "null plan");
    if (K < 1 || K > plan->d.nlayers) return fail(SHT_ERR_DIM, "DimensionMismatch: K outside 1..nlayers");
    DeviceGuard guard(plan->d.device);
    const TuneKey key{plan->d.grid, plan->d.nlat_half, plan->d.lmax, plan->d.mmax, plan->d.nlayers,
                      plan->d.precision, K};
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end() && apply_choice(plan, it->second) == SHT_OK) {
            plan->tuned = 1;
            return SHT_OK;
        }
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    void *coeffs = nullptr, *field = nullptr;
    const size_t cb = (size_t)plan->g.nharm * K * 8, fb = (size_t)plan->g.npoints * K * 4;
    if (hipMalloc(&coeffs, cb) != hipSuccess || hipMalloc(&field, fb) != hipSuccess) {
        if (coeffs) (void)hipFree(coeffs);
        return fail(SHT_ERR_ALLOC, "autotune scratch allocation failed");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SHT_OK;
    if (hipMemsetAsync(coeffs, 0, cb, st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        rc = fail(SHT_ERR_HIP, "autotune set-up failed");
    // best of three means of `reps` round trips, after two untimed ones
    auto round_trip_ms = [&](float &best) -> int {
        const int reps = 10;
        best = 1e30f;
        for (int trial = 0; trial < 3; ++trial) {
            for (int i = 0; i < reps + 2; ++i) {
                if (i == 2) (void)hipEventRecord(e0, st);
                int r = sht_synthesis(plan, coeffs, plan->g.nharm, field, plan->g.npoints, K, 0, stream);
                if (!r) r = sht_analysis(plan, field, plan->g.npoints, coeffs, plan->g.nharm, K, stream);
                if (r) return r;
            }
            (void)hipEventRecord(e1, st);
            float ms = 0.f;
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                return fail(SHT_ERR_HIP, "autotune timing failed");
            best = std::min(best, ms / reps);
        }
        return SHT_OK;
    };
    TuneChoice best{policy_of(plan->pol), plan->leg_xcd, plan->fft_pair, plan->fourier, plan->dft_cg};
    if (!rc) {
        // Fourier stage: each factorisation policy of the LDS FFT, then the ring DFT on MFMA
        // (FP32, and FP16 hi/lo on split plans) with 4 or 8 columns per workgroup
        float best_ms = 1e30f;
        std::vector<TuneChoice> cands;
        for (int q = 0; q < kPolicies; ++q)
            if (plan->pol_ok[q]) cands.push_back({policy_of(q), best.leg_xcd, best.fft_pair, sht::kFourierRadix, 4});
        if (kAutotuneDft)
            for (int kind : {sht::kFourierDft32, sht::kFourierDft16})
                for (int cg : {4, 8})
                    if (kind == sht::kFourierDft32 || plan->d.precision != SHT_PREC_FP32)
                        cands.push_back({best.pol, best.leg_xcd, best.fft_pair, kind, cg});
        for (const TuneChoice &c : cands) {
            float ms = 0.f;
            if ((rc = apply_choice(plan, c)) || (rc = round_trip_ms(ms))) break;
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
        if (!rc) rc = apply_choice(plan, best);
        best_ms = 1e30f;
        TuneChoice b2 = best;
        for (int xcd = 3; xcd >= 0 && !rc; --xcd)
            for (int pair = 1; pair >= 0 && !rc; --pair) {
                if (pair == 0 && best.fourier != sht::kFourierRadix) continue;  // the DFT has one order
                TuneChoice c = best;
                c.leg_xcd = xcd;
                c.fft_pair = pair;
                float ms = 0.f;
                if ((rc = apply_choice(plan, c)) || (rc = round_trip_ms(ms))) break;
                if (ms < best_ms) {
                    best_ms = ms;
                    b2 = c;
                }
            }
        best = b2;
        if (!rc) rc = apply_choice(plan, best);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(st);
    (void)hipFree(coeffs);
    (void)hipFree(field);
    if (rc) return rc;
    plan->tuned = 1;
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tune_cache[key] = best;
    return SHT_OK;
}

int sht_plan_set_graphs(sht_plan *plan, int32_t enable, int32_t flags) {
    if (!plan || enable < 0 || enable > 1 || (flags & ~SHT_GRAPHS_FORCE_FAILURE)) return fail(SHT_ERR_ARG, "bad arguments");
    std::lock_guard<std::mutex> lock(plan->graph_mutex);
    plan->graphs = enable != 0;
    plan->graph_force_failure = (flags & SHT_GRAPHS_FORCE_FAILURE) != 0;
    return SHT_OK;
}

int sht_plan_graph_count(sht_plan *plan, int32_t *cached, int32_t *failed) {
    if (!plan) return fail(SHT_ERR_ARG, "null plan");
    std::lock_guard<std::mutex> lock(plan->graph_mutex);
    int n = 0;
    for (auto &gc : plan->graph_cache)
        for (auto &kv : gc) n += kv.second ? 1 : 0;
    if (cached) *cached = n;
    if (failed) *failed = plan->graph_failed;
    return SHT_OK;
}

int sht_plan_set_tuning(sht_plan *plan, const int32_t *in) {
    if (!plan || !in) return fail(SHT_ERR_ARG, "bad arguments");
    if ((in[0] != 11 && in[0] != 17) || in[1] < 0 || in[1] > 2 || in[2] < 0 || in[2] > 1 || in[3] < 0 ||
        in[3] > 3 || in[4] < 0 || in[4] > 1 || in[5] < 0 || in[5] > 2 || (in[6] != 4 && in[6] != 8))
        return fail(SHT_ERR_ARG, "tuning values out of range");
    FFTPolicy pol;
    pol.mfma_min = in[0];
    pol.eights = in[1];
    pol.odd_first = in[2] != 0;
    return apply_choice(plan, TuneChoice{pol, in[3], in[4], in[5], in[6]});
}

int sht_plan_tuning(const sht_plan *plan, int32_t *out) {
    if (!plan || !out) return fail(SHT_ERR_ARG, "bad arguments");
    out[0] = plan->tuned;
    const FFTPolicy pol = policy_of(plan->pol);
    out[1] = pol.mfma_min;
    out[2] = pol.eights;
    out[3] = pol.odd_first ? 1 : 0;
    out[4] = plan->leg_xcd;
    out[5] = plan->fft_pair;
    out[6] = plan->fourier;
    out[7] = plan->dft_cg;
    return SHT_OK;
}

int sht_gauss_legendre(int32_t nlat, double *nodes, double *weights) {
    if (nlat < 1 || !nodes || !weights) return fail(SHT_ERR_ARG, "bad arguments");
    std::vector<double> x, w;
    gauss_legendre(nlat, x, w);
    std::memcpy(nodes, x.data(), sizeof(double) * nlat);
    std::memcpy(weights, w.data(), sizeof(double) * nlat);
    return SHT_OK;
}

int sht_legendre_column(int32_t lmax, int32_t mmax, double x, double *out) {
    if (mmax < 1 || lmax < mmax || !out) return fail(SHT_ERR_ARG, "bad arguments");
    const double s = std::sqrt(std::max(0.0, 1.0 - x * x));
    double pmm = 0.0;
    long long off = 0;
    for (int m = 0; m < mmax; ++m) {
        pmm = next_pmm(pmm, m, s);
        legendre_order(lmax, m, x, pmm, out + off);
        off += lmax - m;
    }
    return SHT_OK;
}

}  // extern "C"

// ================================================================== batches

struct sht_batch {
    int device = 0;
    int mode = 0;
    std::vector<void *> allocs;
    const DevPlan *plans = nullptr;
    const CallArgs *calls = nullptr;
    const ColRef *colmap = nullptr;  // merged Legendre column groups, passed as single.in
    const Tile *tiles[4] = {nullptr, nullptr, nullptr, nullptr};
    int ntiles[4] = {0, 0, 0, 0};
    unsigned fft_smem = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = false;
};

namespace {

void free_batch(sht_batch *b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (void *a : b->allocs) (void)hipFree(a);
    delete b;
}

const void *stage_func(const sht_batch *b, int st) {
    return (st == sht::kLegendreSynthesis || st == sht::kLegendreAnalysis) ? sht::legendre_kernel(st, b->mode, true)
                                                                            : sht::fft_kernel(st, true);
}

int build_batch(sht_batch *b, const sht_batch_item *items, int n, int flags) {
    std::vector<sht_plan *> plans;
    std::vector<CallArgs> calls;
    std::vector<LaunchUnit> st_units[4];
    std::map<std::tuple<int, int, int>, std::vector<int>> leg_calls;  // (plan, stage, unscale) -> calls
    std::vector<ColRef> colmap;                                     // merged column groups
    int rc;
    // scratch: one Fourier buffer per item (a round trip's two calls share it)
    std::vector<size_t> fb_off(n);
    size_t fb_total = 0;
    for (int i = 0; i < n; ++i) {
        fb_off[i] = fb_total;
        fb_total += (size_t)items[i].plan->g.fb_rows * items[i].K * 16;
        fb_total = (fb_total + 255) & ~(size_t)255;
    }
    void *scratch = nullptr;
    hipError_t e = hipMalloc(&scratch, std::max<size_t>(fb_total, 256));
    if (e != hipSuccess) return fail(SHT_ERR_ALLOC, std::string("hipMalloc batch scratch: ") + hipGetErrorString(e));
    b->allocs.push_back(scratch);
    for (int i = 0; i < n; ++i) {
        const sht_batch_item &it = items[i];
        sht_plan *pl = it.plan;
        int pi = (int)(std::find(plans.begin(), plans.end(), pl) - plans.begin());
        if (pi == (int)plans.size()) plans.push_back(pl);
        void *fb = static_cast<char *>(scratch) + fb_off[i];
        const int K = it.K;
        const int cgs = (K + sht::kColsComplex - 1) / sht::kColsComplex;
        const int kgs = (K + pl->P.kg - 1) / pl->P.kg;
        auto add_fft = [&](int st, int call) {  // tiles (call, ring rank, column group)
            const std::vector<int> &order = pl->h_ring_order_v[pl->pol];
            for (int r = 0; r < (int)order.size(); ++r)
                for (int q = 0; q < kgs; q += 2) {
                    LaunchUnit u{0.0, {}};
                    for (int h = q; h < std::min(q + 2, kgs); ++h) {
                        u.tiles.push_back({call, r, h, 0});
                        u.cost += pl->h_ring_cost_v[pl->pol][order[r]];
                    }
                    st_units[st].push_back(u);
                }
        };
        // Legendre calls are grouped by plan (and unscale flag): after the loop, the calls of a
        // group share column groups (merged_groups below)
        auto add_leg = [&](int st, int call, const std::vector<Tile> &, const std::vector<double> &) {
            leg_calls[{pi, st, calls[call].unscale}].push_back(call);
        };
        if (it.direction == SHT_DIR_SYNTHESIS || it.direction == SHT_DIR_ROUNDTRIP) {
            void *grid = it.direction == SHT_DIR_ROUNDTRIP ? it.mid : it.out;
            const long long ldg = it.direction == SHT_DIR_ROUNDTRIP ? it.ld_mid : it.ld_out;
            calls.push_back({pi, K, it.unscale_coslat ? 1 : 0, 0, it.in, grid, it.ld_in, ldg, fb});
            const int call = (int)calls.size() - 1;
            add_leg(sht::kLegendreSynthesis, call, pl->h_syn, pl->h_syn_cost);
            add_fft(sht::kFFTSynthesis, call);
        }
        if (it.direction == SHT_DIR_ANALYSIS || it.direction == SHT_DIR_ROUNDTRIP) {
            const void *grid = it.direction == SHT_DIR_ROUNDTRIP ? it.mid : it.in;
            const long long ldg = it.direction == SHT_DIR_ROUNDTRIP ? it.ld_mid : it.ld_in;
            calls.push_back({pi, K, 0, 0, grid, it.out, ldg, it.ld_out, fb});
            const int call = (int)calls.size() - 1;
            add_fft(sht::kFFTAnalysis, call);
            add_leg(sht::kLegendreAnalysis, call, pl->h_ana, pl->h_ana_cost);
        }
    }
    // Legendre units.  The calls of one (plan, stage, unscale) group fill column groups of
    // kColsComplex columns in call order, so a workgroup (order m, ring or degree block) reads
    // Λ once for the columns of several items; a group of one call keeps its own column
    // groups.  Units gather the tiles of one order m of a group, so they run back to back
    // on one XCD.
    for (const auto &kv : leg_calls) {
        const int pi = std::get<0>(kv.first), st = std::get<1>(kv.first);
        const sht_plan *pl = plans[pi];
        const std::vector<Tile> &t = st == sht::kLegendreSynthesis ? pl->h_syn : pl->h_ana;
        const std::vector<double> &c = st == sht::kLegendreSynthesis ? pl->h_syn_cost : pl->h_ana_cost;
        const std::vector<int> &cl = kv.second;
        if (cl.size() == 1) {
            const int call = cl[0], cgs = (calls[call].K + sht::kColsComplex - 1) / sht::kColsComplex;
            for (const LaunchUnit &u : legendre_units(t, c, call, cgs)) st_units[st].push_back(u);
            continue;
        }
        const int g0 = (int)(colmap.size() / sht::kColsComplex);
        for (int call : cl)
            for (int k = 0; k < calls[call].K; ++k) colmap.push_back({call, k});
        while (colmap.size() % sht::kColsComplex) colmap.push_back({-1, 0});
        const int ng = (int)(colmap.size() / sht::kColsComplex) - g0;
        std::vector<LaunchUnit> units;
        std::vector<int> unit_of_m;
        for (size_t q = 0; q < t.size(); ++q) {
            const int m = t[q].m;
            if (m >= (int)unit_of_m.size()) unit_of_m.resize(m + 1, -1);
            if (unit_of_m[m] < 0) {
                unit_of_m[m] = (int)units.size();
                units.push_back({0.0, {}});
            }
            LaunchUnit &u = units[unit_of_m[m]];
            for (int g = 0; g < ng; ++g) {
                u.tiles.push_back({-1 - (g0 + g), m, t[q].blk, 0});
                u.cost += c[q];
            }
        }
        st_units[st].insert(st_units[st].end(), units.begin(), units.end());
    }
    if (colmap.empty()) colmap.push_back({-1, 0});
    if ((rc = dev_upload(b->allocs, colmap, &b->colmap))) return rc;
    std::vector<DevPlan> dplans;
    for (sht_plan *pl : plans) {
        dplans.push_back(pl->P);
        b->fft_smem = std::max<unsigned>(b->fft_smem, (unsigned)pl->P.fft_smem);
    }
    if ((rc = dev_upload(b->allocs, dplans, &b->plans))) return rc;
    if ((rc = dev_upload(b->allocs, calls, &b->calls))) return rc;
    for (int st = 0; st < 4; ++st) {
        // heaviest units first across all items, spread over the XCDs
        const std::vector<Tile> order = st_units[st].empty() ? std::vector<Tile>() : xcd_order(st_units[st]);
        b->ntiles[st] = (int)order.size();
        if (!order.empty() && (rc = dev_upload(b->allocs, order, &b->tiles[st]))) return rc;
    }
    if (b->fft_smem > 65536) {
        for (int st : {sht::kFFTSynthesis, sht::kFFTAnalysis})
            (void)hipFuncSetAttribute(sht::fft_kernel(st, true), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)b->fft_smem);
    }
    if (flags & SHT_BATCH_NO_GRAPH) return SHT_OK;
    // explicit graph: one kernel node per non-empty stage, chained in stage order
    if (hipGraphCreate(&b->graph, 0) != hipSuccess) return SHT_OK;
    hipGraphNode_t prev = nullptr;
    CallArgs none{};
    none.in = b->colmap;
    for (int st = 0; st < 4; ++st) {
        if (b->ntiles[st] == 0) continue;
        const DevPlan *a0 = b->plans;
        CallArgs a1 = none;
        const CallArgs *a2 = b->calls;
        const Tile *a3 = b->tiles[st];
        void *args[] = {&a0, &a1, &a2, &a3};
        hipKernelNodeParams kp{};
        kp.func = const_cast<void *>(stage_func(b, st));
        kp.gridDim = dim3(b->ntiles[st]);
        kp.blockDim = dim3(256);
        kp.sharedMemBytes = (st == sht::kFFTSynthesis || st == sht::kFFTAnalysis) ? b->fft_smem : 0;
        kp.kernelParams = args;
        kp.extra = nullptr;
        hipGraphNode_t node;
        if (hipGraphAddKernelNode(&node, b->graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp) != hipSuccess) {
            (void)hipGetLastError();
            return SHT_OK;  // direct launches
        }
        prev = node;
    }
    if (flags & SHT_BATCH_FORCE_GRAPH_FAILURE) return SHT_OK;
    if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        b->exec = nullptr;
        return SHT_OK;
    }
    b->use_graph = true;
    return SHT_OK;
}

}  // namespace

extern "C" {

int sht_batch_create(const sht_batch_item *items, int32_t n, int32_t flags, sht_batch **batch) {
    sht::TraceRange trace_range("sht_batch_create");
    if (!batch) return fail(SHT_ERR_ARG, "null batch pointer");
    *batch = nullptr;
    if (!items || n < 1) return fail(SHT_ERR_ARG, "empty batch");
    for (int i = 0; i < n; ++i) {
        const sht_batch_item &it = items[i];
        if (!it.plan) return fail(SHT_ERR_ARG, "batch item " + std::to_string(i) + ": null plan");
        if (it.direction < SHT_DIR_SYNTHESIS || it.direction > SHT_DIR_ROUNDTRIP)
            return fail(SHT_ERR_ARG, "batch item " + std::to_string(i) + ": bad direction");
        if (it.plan->d.device != items[0].plan->d.device || it.plan->d.precision != items[0].plan->d.precision)
            return fail(SHT_ERR_UNSUPPORTED, "batch plans must share device and precision");
        const int rc = check_call(it.plan, it.in, it.out, it.K);
        if (rc) return rc;
        if (it.direction == SHT_DIR_ROUNDTRIP && !it.mid) return fail(SHT_ERR_ARG, "round trip needs a grid buffer");
        const long long nh = it.plan->g.nharm, np = it.plan->g.npoints;
        const bool spec_in = it.direction != SHT_DIR_ANALYSIS;
        if (it.ld_in < (spec_in ? nh : np) || it.ld_out < (it.direction == SHT_DIR_SYNTHESIS ? np : nh) ||
            (it.direction == SHT_DIR_ROUNDTRIP && it.ld_mid < np))
            return fail(SHT_ERR_DIM, "DimensionMismatch: batch item " + std::to_string(i) + " column stride too small");
    }
    sht_batch *b = new (std::nothrow) sht_batch();
    if (!b) return fail(SHT_ERR_ALLOC, "host allocation failed");
    b->device = items[0].plan->d.device;
    b->mode = items[0].plan->d.precision;
    DeviceGuard guard(b->device);
    const int rc = build_batch(b, items, n, flags);
    if (rc) {
        free_batch(b);
        return rc;
    }
    *batch = b;
    return SHT_OK;
}

int sht_batch_launch(sht_batch *b, void *stream) {
    sht::TraceRange trace_range("sht_batch_launch");
    if (!b) return fail(SHT_ERR_ARG, "null batch");
    DeviceGuard guard(b->device);
    hipError_t e;
    if (b->use_graph) {
        e = hipGraphLaunch(b->exec, reinterpret_cast<hipStream_t>(stream));
        if (e != hipSuccess) return fail(SHT_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
        return SHT_OK;
    }
    CallArgs none{};
    none.in = b->colmap;
    for (int st = 0; st < 4; ++st) {
        if (b->ntiles[st] == 0) continue;
        const unsigned smem = (st == sht::kFFTSynthesis || st == sht::kFFTAnalysis) ? b->fft_smem : 0;
        e = launch_stage(stage_func(b, st), dim3(b->ntiles[st]), smem, b->plans, none, b->calls, b->tiles[st], stream);
        if (e != hipSuccess) return fail(SHT_ERR_HIP, std::string("batch stage launch: ") + hipGetErrorString(e));
    }
    return SHT_OK;
}

int sht_batch_uses_graph(const sht_batch *b) { return b && b->use_graph ? 1 : 0; }

int sht_batch_destroy(sht_batch *b) {
    if (!b) return SHT_OK;
    DeviceGuard guard(b->device);
    (void)hipDeviceSynchronize();
    free_batch(b);
    return SHT_OK;
}

}  // extern "C"
