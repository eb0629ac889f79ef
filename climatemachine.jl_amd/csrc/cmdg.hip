// cmdg.hip -- MI355X (gfx950) DG right-hand side and LSRK54 stepper for the dry
// AtmosModel balance law, behind the C ABI declared in include/cmdg.h.
//
// One workgroup owns one hexahedral element (Np = Nq^3 nodes, one lane per node,
// rounded up to whole 64-lane wavefronts).  Per right-hand side:
//
//   k_gradient   (diffusive models only) gradient argument G(q) staged in LDS,
//                tensor-product derivative along r/s/t, lifted central face jumps
//                evaluated by the lanes that own the face nodes, gradient-flux
//                transform; writes the gradient flux (Np, ngf, nelem).
//   k_tendency   first + second order flux at each node, contravariant flux
//                M (xi_a . F) staged in LDS, weak divergence D^T along r/s/t,
//                gravity source, Rusanov/central face fluxes evaluated by the
//                lanes owning the face nodes (periodic neighbours come from vmap+
//                on the device), and an epilogue that either writes
//                alpha*RHS + beta*tendency or performs the whole low-storage
//                Runge-Kutta stage update.
//
// Reference kernels replaced: volume_tendency! (DGModel_kernels.jl:64-548),
// dgsem_interface_tendency! (:588-901), volume_gradients! (:934-1328),
// dgsem_interface_gradients! (:1365-1651), kernel_nodal_update_auxiliary_state!
// (:1769-1825) and the LSRK update! kernel (LowStorageRungeKuttaMethod.jl:146-158).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

#include "../../include/cmdg.h"

#include "cmdg_physics.h"
#include "cmdg_general.h"
#include "cmdg_affine.h"
#include "cmdg_mixed.h"
#include "cmdg_learn.h"
#include "cmdg_diag.h"

namespace {
using namespace cmdg;

// update! (LowStorageRungeKuttaMethod.jl:146-158)
__global__ void k_lsrk_update(double *__restrict__ dQ, double *__restrict__ Q, int64_t n, double rka, double bdt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const double d = dQ[i];
        Q[i] += bdt * d;
        dQ[i] = d * rka;
    }
}

__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// LSRK54CarpenterKennedy coefficients (LowStorageRungeKuttaMethod.jl:302-324)
const double RKA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                       -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
const double RKB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                       1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                       2277821191437.0 / 14882151754819.0};

}  // namespace

// ====================================================================== context
struct cmdg_ctx {
    int N, Nq, Np, Nfp;
    int64_t nelem;
    int ngf;
    bool diffusive, smag;
    KParams P;
    KArrays A;
    hipStream_t stream;
    double *d_vg = nullptr, *d_sg = nullptr, *d_D = nullptr, *d_aux = nullptr, *d_gf = nullptr, *d_Qs = nullptr;
    int32_t *d_nbr = nullptr, *d_bnd = nullptr;
    // affine-brick specialisation (see cmdg_affine.h)
    bool affine = false;
    AffArrays AA{};
    double *d_geo = nullptr, *d_W = nullptr, *d_Wf = nullptr;
    int32_t *d_nbe = nullptr;
    double *d_f2tr = nullptr, *d_str = nullptr;  // face traces
    int32_t *d_cls = nullptr;  // static-field class of each element (null: none)
    double *d_sa = nullptr;    // static fields by class
    int32_t *d_order = nullptr;  // tiled traversal order of the affine launches (null: none)
    const int32_t *ord_grad = nullptr, *ord_tend = nullptr;  // per pass: d_order or null (storage)
    // mixed precision (FP32 fluxes and gradient flux)
    bool mixed = false;
    MixParams MP{};
    // the FP32 gradient flux in stress form: iso32 its isotropic layout (no buoyancy
    // correction of the viscosity), ngf32 its planes
    float *d_gf32 = nullptr, *d_gphi32 = nullptr;
    bool iso32 = true;
    int ngf32 = 0;
    // learned closure (mixed precision only)
    bool learned = false, mlp_ready = false;
    MlpDev *d_mlp = nullptr;
    // diagnostics (Courant numbers, weighted sums): reduction scratch, allocated on first use
    int32_t aux_coord = -1;
    double2 *d_part = nullptr;
    double *d_scalar = nullptr;
    int64_t npart = 0;
    std::string err;
    // optional per-kernel event timing (cmdg_set_kernel_timing)
    bool timing = false;
    struct Rec {
        int kind;
        hipEvent_t start, stop;
    };
    std::vector<Rec> recs;
};

namespace {

int fail(cmdg_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

int check_hip(cmdg_ctx *c, hipError_t e, const char *what) {
    if (e == hipSuccess) return CMDG_OK;
    return fail(c, e == hipErrorOutOfMemory ? CMDG_ENOMEM : CMDG_EHIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

// kernel kinds reported by cmdg_kernel_timing
enum { KIND_GRADIENT = 0, KIND_TENDENCY = 1, KIND_UPDATE = 2, KIND_COPY = 3, NKIND = 4 };

struct TimedLaunch {
    cmdg_ctx *c;
    int kind;
    hipEvent_t stop = nullptr;
    TimedLaunch(cmdg_ctx *c_, int kind_) : c(c_), kind(kind_) {
        if (!c->timing) return;
        hipEvent_t start;
        if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess) { c->timing = false; return; }
        (void)hipEventRecord(start, c->stream);
        c->recs.push_back({kind, start, stop});
    }
    ~TimedLaunch() {
        if (stop) (void)hipEventRecord(stop, c->stream);
    }
};

template <class T> int upload(cmdg_ctx *c, T **dst, const T *src, size_t count) {
    hipError_t e = hipMalloc((void **)dst, count * sizeof(T));
    if (e != hipSuccess) return check_hip(c, e, "hipMalloc");
    if (src) {
        e = hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return check_hip(c, e, "hipMemcpy");
    }
    return CMDG_OK;
}

// ES: the elements of this launch (affine kernels; the general ones take them all).
template <int Nq>
int launch_gradient_impl(cmdg_ctx *c, const double *Q, double *Qcopy, int write_aux, const ElemSet &ES) {
    constexpr int NT = threads_for<Nq>();
    dim3 grid((unsigned)xcd_grid(c->affine ? ES.count : c->nelem)), block(NT);
    TimedLaunch tl(c, KIND_GRADIENT);
    // N^2 (Smagorinsky-Lilly) needs theta_v only under a gravitational potential
    const bool buoy = c->P.a_gPhi >= 0;
#define CMDG_GRAD_MIX(SMAG, LEARN, BUOY)                                                                       \
    hipLaunchKernelGGL((k_gradient_mix<Nq, SMAG, LEARN, BUOY>), grid, block, 0, c->stream, c->P, c->MP, c->AA, ES, \
                       Q, c->d_aux, c->d_gf32, write_aux, c->d_mlp)
#define CMDG_GRAD_AFF(SMAG, BUOY)                                                                               \
    hipLaunchKernelGGL((k_gradient_aff<Nq, SMAG, BUOY>), grid, block, 0, c->stream, c->P, c->AA, ES, Q, c->d_aux, \
                       c->d_gf, c->d_f2tr, write_aux)
    if (c->mixed) {
        if (c->learned) {
            if (buoy) CMDG_GRAD_MIX(true, true, true);
            else CMDG_GRAD_MIX(true, true, false);
        } else if (c->smag) {
            if (buoy) CMDG_GRAD_MIX(true, false, true);
            else CMDG_GRAD_MIX(true, false, false);
        } else
            CMDG_GRAD_MIX(false, false, false);
    } else if (c->affine) {
        if (c->smag) {
            if (buoy) CMDG_GRAD_AFF(true, true);
            else CMDG_GRAD_AFF(true, false);
        } else
            CMDG_GRAD_AFF(false, false);
    } else if (c->smag)
        hipLaunchKernelGGL((k_gradient<Nq, true>), grid, block, 0, c->stream, c->P, c->A, Q, c->d_aux, c->d_gf, Qcopy, write_aux);
    else
        hipLaunchKernelGGL((k_gradient<Nq, false>), grid, block, 0, c->stream, c->P, c->A, Q, c->d_aux, c->d_gf, Qcopy, write_aux);
#undef CMDG_GRAD_MIX
#undef CMDG_GRAD_AFF
    return check_hip(c, hipGetLastError(), "gradient launch");
}

template <int Nq>
int launch_tendency_impl(cmdg_ctx *c, const double *Q, double *tend, double alpha, double beta, int mode,
                         double *Qout, double rka, double bdt, int write_aux, int dq_store, const ElemSet &ES) {
    constexpr int NT = threads_for<Nq>();
    dim3 grid((unsigned)xcd_grid(c->affine ? ES.count : c->nelem)), block(NT);
    TimedLaunch tl(c, KIND_TENDENCY);
#define CMDG_TEND_MIX(DIFF, ISO, MODE)                                                                           \
    hipLaunchKernelGGL((k_tendency_mix<Nq, DIFF, ISO, MODE>), grid, block, 0, c->stream, c->P, c->MP, c->AA, ES, Q, \
                       c->d_gf32, c->d_aux, tend, alpha, beta, Qout, rka, bdt, write_aux, dq_store)
#define CMDG_TEND(DIFF, SMAG, MODE)                                                                               \
    do {                                                                                                          \
        if (c->mixed) {                                                                                           \
            if (!DIFF || c->iso32) CMDG_TEND_MIX(DIFF, true, MODE);                                               \
            else CMDG_TEND_MIX(DIFF, false, MODE);                                                                \
        } else if (c->affine)                                                                                     \
            hipLaunchKernelGGL((k_tendency_aff<Nq, DIFF, SMAG, MODE>), grid, block, 0, c->stream, c->P, c->AA,    \
                               ES, Q, c->d_gf, c->d_f2tr, c->d_aux, tend, alpha, beta, Qout, rka, bdt,      \
                               write_aux, dq_store);                                                              \
        else                                                                                                      \
            hipLaunchKernelGGL((k_tendency<Nq, DIFF, SMAG, MODE>), grid, block, 0, c->stream, c->P, c->A, Q,      \
                               c->d_gf, c->d_aux, tend, alpha, beta, Qout, rka, bdt, write_aux, dq_store);       \
    } while (0)
    if (mode == 0) {
        if (!c->diffusive) CMDG_TEND(false, false, 0);
        else if (c->smag) CMDG_TEND(true, true, 0);
        else CMDG_TEND(true, false, 0);
    } else {
        if (!c->diffusive) CMDG_TEND(false, false, 1);
        else if (c->smag) CMDG_TEND(true, true, 1);
        else CMDG_TEND(true, false, 1);
    }
#undef CMDG_TEND
#undef CMDG_TEND_MIX
    return check_hip(c, hipGetLastError(), "tendency launch");
}

#define CMDG_DISPATCH_NQ(CALL)                                                 \
    switch (c->Nq) {                                                           \
    case 4: return CALL(4);                                                    \
    case 5: return CALL(5);                                                    \
    default: return fail(c, CMDG_EUNSUPPORTED, "polynomial order not compiled"); \
    }

int launch_gradient(cmdg_ctx *c, const double *Q, double *Qcopy, int write_aux, const ElemSet *ES = nullptr) {
    const ElemSet all{c->ord_grad, 0, c->nelem};
#define CALL(NQ) launch_gradient_impl<NQ>(c, Q, Qcopy, write_aux, ES ? *ES : all)
    CMDG_DISPATCH_NQ(CALL)
#undef CALL
}

// mode 1 (LSRK stage): beta 1 reads dQ, beta 0 takes it as zeros without reading it;
// dq_store 0 leaves the stage's scaled dQ unstored (see k_tendency_mix)
int launch_tendency(cmdg_ctx *c, const double *Q, double *tend, double alpha, double beta, int mode, double *Qout,
                    double rka, double bdt, int write_aux, int dq_store = 1, const ElemSet *ES = nullptr) {
    const ElemSet all{c->ord_tend, 0, c->nelem};
#define CALL(NQ)                                                                                                 \
    launch_tendency_impl<NQ>(c, Q, tend, alpha, beta, mode, Qout, rka, bdt, write_aux, dq_store, ES ? *ES : all)
    CMDG_DISPATCH_NQ(CALL)
#undef CALL
}

// dQ/dt into tend (mode 0) with the auxiliary refresh of DGModel.jl:110-116
int launch_rhs(cmdg_ctx *c, const double *Q, double *tend, double alpha, double beta) {
    int rc;
    if (c->learned && !c->mlp_ready) return fail(c, CMDG_EINVAL, "learned closure weights not set");
    if (c->diffusive && (rc = launch_gradient(c, Q, nullptr, 1))) return rc;
    return launch_tendency(c, Q, tend, alpha, beta, 0, nullptr, 0.0, 0.0, c->diffusive ? 0 : 1);
}

int launch_copy(cmdg_ctx *c, double *dst, const double *src, int64_t n) {
    TimedLaunch tl(c, KIND_COPY);
    hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, c->stream, dst, src, n);
    return check_hip(c, hipGetLastError(), "copy launch");
}

// Exact scale c with c * base[i] == vals[i] for all i (NaN if none near vals[0] / base[0]).
double exact_scale(const double *vals, int64_t vstride, const double *base, int count) {
    double c0 = vals[0] / base[0];
    double cands[5] = {c0, std::nextafter(c0, INFINITY), std::nextafter(c0, -INFINITY), 0, 0};
    cands[3] = std::nextafter(cands[1], INFINITY);
    cands[4] = std::nextafter(cands[2], -INFINITY);
    for (double c : cands) {
        bool ok = true;
        for (int i = 0; i < count && ok; ++i) ok = (c * base[i] == vals[i * vstride]);
        if (ok) return c;
    }
    return NAN;
}

int fmask_vid(int Nq, int f, int fn) {
    const int a = fn % Nq, b = fn / Nq, N = Nq - 1;
    switch (f) {
    case 0: return 0 + Nq * (a + Nq * b);
    case 1: return N + Nq * (a + Nq * b);
    case 2: return a + Nq * (0 + Nq * b);
    case 3: return a + Nq * (N + Nq * b);
    case 4: return a + Nq * (b + Nq * 0);
    default: return a + Nq * (b + Nq * N);
    }
}

// Classes of elements whose static auxiliary fields (Phi, grad Phi, rho_ref, p_ref,
// Delta: what the affine kernels read through AffArrays::sa) agree bit for bit.  On a
// brick whose reference state depends on height alone there is one class per element
// layer.  Returns false (no classes) unless they are at least 4 times fewer than the
// elements; rep[c] is the first element of class c.
bool static_classes(const cmdg_desc *d, const KParams &P, int Np, std::vector<int32_t> &cls,
                    std::vector<int64_t> &rep) {
    const int64_t E = d->nelem;
    int fields[7], nf = 0;
    for (int f : {P.a_Phi, P.a_gPhi, P.a_gPhi >= 0 ? P.a_gPhi + 1 : -1, P.a_gPhi >= 0 ? P.a_gPhi + 2 : -1,
                  P.a_ref_rho, P.a_ref_p, P.a_Delta})
        if (f >= 0) fields[nf++] = f;
    cls.assign(E, 0);
    rep.clear();
    if (nf == 0) {
        rep.push_back(0);
        return E >= 4;
    }
    auto key = [&](int64_t e, std::string &k) {
        k.resize(sizeof(double) * Np * nf);
        for (int i = 0; i < nf; ++i)
            std::memcpy(&k[sizeof(double) * Np * i], d->aux_init + (int64_t)Np * ((int64_t)d->naux * e + fields[i]),
                        sizeof(double) * Np);
    };
    std::unordered_map<size_t, std::vector<int32_t>> byhash;
    std::string k, kr;
    for (int64_t e = 0; e < E; ++e) {
        key(e, k);
        std::vector<int32_t> &bucket = byhash[std::hash<std::string_view>{}(k)];
        int32_t found = -1;
        for (int32_t c : bucket) {
            key(rep[c], kr);
            if (kr == k) {
                found = c;
                break;
            }
        }
        if (found < 0) {
            found = (int32_t)rep.size();
            rep.push_back(e);
            bucket.push_back(found);
            if ((int64_t)rep.size() * 4 > E) return false;
        }
        cls[e] = found;
    }
    return true;
}

// Traversal order of the affine kernels' launches.  A block gathers its face
// neighbours' state from their volume; when the neighbour's block ran shortly before
// on the same XCD those lines are still in that XCD's L2.  The storage order of a
// stacked brick puts the y neighbours a whole x-y row of columns apart (625 elements
// at 25^3, 9216 at 96^3: tens of MB, far beyond the 4 MB L2), so launches walk the
// elements in T x T x T tiles instead: z fastest inside a tile, then x, then y, and
// the tiles in the same order -- every neighbour inside a tile is at most T^2
// positions away.  The lattice position of each element comes from its centroid
// (vgeo x1, x2, x3); returns false, keeping the storage order, when the centroids do
// not form a complete lattice or the descriptor has no coordinates.
bool tiled_order_lattice(const std::vector<int64_t> idx[3], const int64_t nax[3], int64_t E, const int T[3],
                         std::vector<int32_t> &order);

bool tiled_order(const cmdg_desc *d, int Np, const int T[3], std::vector<int32_t> &order) {
    const int64_t E = d->nelem;
    if (T[0] < 0 || T[1] < 0 || T[2] < 0 || d->nvgeo < 15 || E > INT32_MAX) return false;
    std::vector<int64_t> idx[3];
    int64_t nax[3];
    for (int a = 0; a < 3; ++a) {
        std::vector<double> c(E);
        for (int64_t e = 0; e < E; ++e) {
            const double *x = d->vgeo + (int64_t)Np * ((int64_t)d->nvgeo * e + 12 + a);
            double sum = 0.0;
            for (int n = 0; n < Np; ++n) sum += x[n];
            c[e] = sum / Np;
        }
        std::vector<double> u(c);
        std::sort(u.begin(), u.end());
        const double tol = 1e-9 * (std::fabs(u.front()) + std::fabs(u.back()) + 1e-300);
        std::vector<double> lev;  // one value per lattice plane
        for (double v : u)
            if (lev.empty() || v - lev.back() > tol) lev.push_back(v);
        nax[a] = (int64_t)lev.size();
        idx[a].resize(E);
        for (int64_t e = 0; e < E; ++e) {
            auto it = std::lower_bound(lev.begin(), lev.end(), c[e] - tol);
            if (it == lev.end() || std::fabs(*it - c[e]) > tol) return false;
            idx[a][e] = it - lev.begin();
        }
    }
    return tiled_order_lattice(idx, nax, E, T, order);
}

// The tile order of elements at lattice positions idx[a][e] of an nax[0] x nax[1] x
// nax[2] lattice, tiles of T[0] x T[1] x T[2] elements (0: the whole axis), z fastest
// then x then y within a tile and between tiles (false when the positions do not fill
// the lattice once each).
bool tiled_order_lattice(const std::vector<int64_t> idx[3], const int64_t nax[3], int64_t E, const int T[3],
                         std::vector<int32_t> &order) {
    if (E > INT32_MAX || nax[0] * nax[1] * nax[2] != E) return false;
    int64_t t[3];
    for (int a = 0; a < 3; ++a) {
        if (T[a] < 0) return false;
        t[a] = T[a] == 0 ? nax[a] : std::min<int64_t>(T[a], nax[a]);
    }
    const int64_t ntx = (nax[0] + t[0] - 1) / t[0], ntz = (nax[2] + t[2] - 1) / t[2];
    std::vector<std::pair<int64_t, int32_t>> key(E);
    std::vector<char> seen(E, 0);
    for (int64_t e = 0; e < E; ++e) {
        const int64_t ix = idx[0][e], iy = idx[1][e], iz = idx[2][e];
        const int64_t cell = iz + nax[2] * (ix + nax[0] * iy);
        if (seen[cell]) return false;  // two elements at one lattice point
        seen[cell] = 1;
        const int64_t tile = iz / t[2] + ntz * (ix / t[0] + ntx * (iy / t[1]));
        key[e] = {((tile * t[1] + iy % t[1]) * t[0] + ix % t[0]) * t[2] + iz % t[2], (int32_t)e};
    }
    std::sort(key.begin(), key.end());
    order.resize(E);
    for (int64_t p = 0; p < E; ++p) order[p] = key[p].second;
    return true;
}

// Bitwise test that the mesh consists of axis-aligned affine bricks with conforming
// neighbours; fills the per-element record used by the affine kernels.
bool detect_affine(const cmdg_desc *d, int Nq, std::vector<double> &geo, std::vector<int32_t> &nbe,
                   std::vector<double> &W, std::vector<double> &Wf) {
    const int Np = Nq * Nq * Nq, Nfp = Nq * Nq;
    const int64_t E = d->nelem;
    if (!d->omega) return false;
    const double *w = d->omega;
    W.resize(Np);
    Wf.resize(Nfp);
    for (int n = 0; n < Np; ++n) W[n] = (w[n / (Nq * Nq)] * w[(n / Nq) % Nq]) * w[n % Nq];
    for (int fn = 0; fn < Nfp; ++fn) Wf[fn] = w[fn / Nq] * w[fn % Nq];
    geo.assign((size_t)AFF_GEO * E, 0.0);
    nbe.assign((size_t)6 * E, -1);
    for (int64_t e = 0; e < E; ++e) {
        const double *v = d->vgeo + (int64_t)Np * d->nvgeo * e;
        double *g = &geo[(size_t)AFF_GEO * e];
        for (int f = 0; f < 9; ++f) {
            const bool diag = (f == 0 || f == 4 || f == 8);
            const double first = v[(int64_t)Np * f];
            for (int n = 0; n < Np; ++n) {
                const double val = v[(int64_t)Np * f + n];
                if (diag ? !(val == first && val != 0.0 && std::isfinite(val)) : val != 0.0) return false;
            }
            if (diag) g[f / 4] = first;
        }
        const double *M = v + (int64_t)Np * 9, *MI = v + (int64_t)Np * 10;
        const double J = exact_scale(M, 1, W.data(), Np);
        if (!(J > 0.0)) return false;
        for (int n = 0; n < Np; ++n)
            if (MI[n] != 1.0 / M[n]) return false;
        g[3] = J;
        for (int f = 0; f < 6; ++f) {
            const double *sgp = d->sgeo + (int64_t)5 * Nfp * (f + 6 * e);
            const double nd = (f & 1) ? 1.0 : -1.0;
            for (int fn = 0; fn < Nfp; ++fn) {
                for (int c = 0; c < 3; ++c)
                    if (sgp[5 * fn + c] != (c == f / 2 ? nd : 0.0)) return false;
                if (sgp[5 * fn + 4] != MI[fmask_vid(Nq, f, fn)]) return false;
            }
            const double sJ = exact_scale(sgp + 3, 5, Wf.data(), Nfp);
            if (!(sJ > 0.0)) return false;
            g[4 + f] = sJ;
            if (d->elemtobndy[f + 6 * e] != 0) continue;
            const int64_t *vp = d->vmap_plus + (int64_t)Nfp * (f + 6 * e);
            const int64_t e2 = (vp[0] - 1) / Np;
            for (int fn = 0; fn < Nfp; ++fn)
                if (vp[fn] - 1 != (int64_t)Np * e2 + fmask_vid(Nq, f ^ 1, fn)) return false;
            nbe[6 * e + f] = (int32_t)e2;
        }
        if (d->aux_Delta >= 0) {
            const double *Dl = d->aux_init + (int64_t)Np * (d->aux_Delta + (int64_t)d->naux * e);
            for (int n = 1; n < Np; ++n)
                if (Dl[n] != Dl[0]) return false;
            g[10] = Dl[0];
        }
    }
    return true;
}

// The per-element records of detect_affine for a brick given by its break points:
// the values the host grid builder puts in vgeo/sgeo (xi = 2/h, J = (h1/2)(h2/2)(h3/2),
// face Jacobians (h_b/2)(h_c/2)), the neighbours of the stacked element order, and
// Delta from the auxiliary state (constant per element).  Returns an error text, empty
// on success.
std::string brick_geometry(const cmdg_desc *d, const cmdg_brick *b, int Nq, std::vector<double> &geo,
                           std::vector<int32_t> &nbe, std::vector<double> &W, std::vector<double> &Wf,
                           std::vector<int64_t> idx[3]) {
    const int Np = Nq * Nq * Nq, Nfp = Nq * Nq;
    const int64_t E = d->nelem, nx = b->ne[0], ny = b->ne[1], nz = b->ne[2];
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx * ny * nz != E) return "brick element counts do not match nelem";
    for (int a = 0; a < 3; ++a) {
        if (!b->x[a]) return "missing brick break points";
        for (int64_t m = 0; m < b->ne[a]; ++m)
            if (!(b->x[a][m + 1] > b->x[a][m])) return "brick break points must increase";
        if (!b->periodic[a] && (b->tag[a][0] == 0 || b->tag[a][1] == 0))
            return "a non-periodic brick face needs a non-zero boundary tag";
    }
    const double *w = d->omega;
    W.resize(Np);
    Wf.resize(Nfp);
    for (int n = 0; n < Np; ++n) W[n] = (w[n / (Nq * Nq)] * w[(n / Nq) % Nq]) * w[n % Nq];
    for (int fn = 0; fn < Nfp; ++fn) Wf[fn] = w[fn / Nq] * w[fn % Nq];
    geo.assign((size_t)AFF_GEO * E, 0.0);
    nbe.assign((size_t)6 * E, -1);
    for (int a = 0; a < 3; ++a) idx[a].resize(E);
    for (int64_t e = 0; e < E; ++e) {
        const int64_t p[3] = {(e / nz) % nx, e / (nz * nx), e % nz};
        double h[3];
        for (int a = 0; a < 3; ++a) {
            idx[a][e] = p[a];
            h[a] = b->x[a][p[a] + 1] - b->x[a][p[a]];
        }
        double *g = &geo[(size_t)AFF_GEO * e];
        g[0] = 2.0 / h[0];
        g[1] = 2.0 / h[1];
        g[2] = 2.0 / h[2];
        g[3] = ((h[0] / 2) * (h[1] / 2)) * (h[2] / 2);
        const double sJ[3] = {(h[1] / 2) * (h[2] / 2), (h[0] / 2) * (h[2] / 2), (h[0] / 2) * (h[1] / 2)};
        for (int f = 0; f < 6; ++f) {
            const int a = f / 2;
            g[4 + f] = sJ[a];
            int64_t q[3] = {p[0], p[1], p[2]};
            q[a] += (f & 1) ? 1 : -1;
            if (q[a] < 0 || q[a] >= b->ne[a]) {
                if (!b->periodic[a]) continue;  // a wall: no neighbour
                q[a] = (q[a] + b->ne[a]) % b->ne[a];
            }
            nbe[6 * e + f] = (int32_t)(q[2] + nz * (q[0] + nx * q[1]));
        }
        if (d->aux_Delta >= 0) {
            const double *Dl = d->aux_init + (int64_t)Np * (d->aux_Delta + (int64_t)d->naux * e);
            for (int n = 1; n < Np; ++n)
                if (Dl[n] != Dl[0]) return "Delta must be constant in each element of a brick";
            g[10] = Dl[0];
        }
    }
    return "";
}

int create_impl(const cmdg_desc *d, const cmdg_brick *brick, void *stream, cmdg_ctx **out);

}  // namespace

extern "C" {

const char *cmdg_version(void) { return "cmdg-mi355x 0.1 (gfx950)"; }

const char *cmdg_last_error(cmdg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cmdg_create(const cmdg_desc *d, void *stream, cmdg_ctx **out) { return create_impl(d, nullptr, stream, out); }

int cmdg_create_brick(const cmdg_desc *d, const cmdg_brick *brick, void *stream, cmdg_ctx **out) {
    if (!brick) {
        if (out) *out = nullptr;
        return CMDG_EINVAL;
    }
    return create_impl(d, brick, stream, out);
}

}  // extern "C"

namespace {

int create_impl(const cmdg_desc *d, const cmdg_brick *brick, void *stream, cmdg_ctx **out) {
    if (!out) return CMDG_EINVAL;
    *out = nullptr;
    if (!d) return CMDG_EINVAL;
    cmdg_ctx *c = new (std::nothrow) cmdg_ctx();
    if (!c) return CMDG_ENOMEM;
    *out = c;
    if (d->N != 3 && d->N != 4) return fail(c, CMDG_EUNSUPPORTED, "only N = 3 and N = 4 are compiled");
    if (d->nelem <= 0 || (!brick && d->nvgeo < 11) || d->naux < 0) return fail(c, CMDG_EINVAL, "bad sizes");
    if (!d->D || !(brick || (d->vgeo && d->sgeo && d->vmap_minus && d->vmap_plus && d->elemtobndy)))
        return fail(c, CMDG_EINVAL, "missing geometry or connectivity");
    if (brick && !d->omega) return fail(c, CMDG_EINVAL, "a brick needs the quadrature weights omega");
    if (d->precision != CMDG_FP64 && d->precision != CMDG_MIXED) return fail(c, CMDG_EINVAL, "unknown precision");
    if (d->turbulence == CMDG_LEARNED_CLOSURE && d->precision != CMDG_MIXED)
        return fail(c, CMDG_EUNSUPPORTED, "the learned closure runs in the mixed-precision path");
    if (d->numerical_flux != CMDG_RUSANOV && d->numerical_flux != CMDG_CENTRAL)
        return fail(c, CMDG_EUNSUPPORTED, "numerical flux");
    auto in_aux = [&](int32_t off, int width) { return off < 0 || off + width <= d->naux; };
    if (!in_aux(d->aux_coord, 3) || !in_aux(d->aux_Phi, 1) || !in_aux(d->aux_gradPhi, 3) || !in_aux(d->aux_ref_rho, 1) ||
        !in_aux(d->aux_ref_p, 1) || !in_aux(d->aux_Delta, 1) || !in_aux(d->aux_theta_v, 1) || !in_aux(d->aux_air_T, 1))
        return fail(c, CMDG_EINVAL, "auxiliary offset out of range");
    if (d->naux > 0 && !d->aux_init) return fail(c, CMDG_EINVAL, "missing initial auxiliary state");
    if (d->orientation == CMDG_FLAT_ORIENTATION && (d->aux_Phi < 0 || d->aux_gradPhi < 0))
        return fail(c, CMDG_EINVAL, "FlatOrientation needs Phi and grad Phi in aux");
    if (d->ref_state == CMDG_HYDROSTATIC_STATE && (d->aux_ref_rho < 0 || d->aux_ref_p < 0))
        return fail(c, CMDG_EINVAL, "HydrostaticState needs ref rho and p in aux");
    if ((d->turbulence == CMDG_SMAGORINSKY_LILLY || d->turbulence == CMDG_LEARNED_CLOSURE) &&
        (d->aux_Delta < 0 || d->aux_theta_v < 0))
        return fail(c, CMDG_EINVAL, "Smagorinsky-Lilly and the learned closure need Delta and theta_v in aux");

    c->N = d->N;
    c->Nq = d->N + 1;
    c->Np = c->Nq * c->Nq * c->Nq;
    c->Nfp = c->Nq * c->Nq;
    c->nelem = d->nelem;
    if ((double)c->nelem * c->Np >= 2147483647.0) return fail(c, CMDG_EUNSUPPORTED, "mesh too large for int32 node ids");
    c->learned = d->turbulence == CMDG_LEARNED_CLOSURE;
    c->aux_coord = d->aux_coord;
    c->smag = d->turbulence == CMDG_SMAGORINSKY_LILLY || c->learned;
    c->diffusive = c->smag || d->viscosity != 0.0;
    c->ngf = c->learned ? 11 : (c->smag ? 10 : 9);
    c->stream = (hipStream_t)stream;

    KParams &P = c->P;
    P.R_d = d->R_d; P.cp_d = d->cp_d; P.cv_d = d->cv_d; P.T_0 = d->T_0; P.grav = d->grav; P.MSLP = d->MSLP;
    P.visc = d->viscosity; P.C_smag = d->C_smag; P.inv_Pr = d->inv_Pr_turb;
    P.inv_cv = 1.0 / d->cv_d;
    P.gammaR = (d->cp_d / d->cv_d) * d->R_d;
    P.kappa = d->R_d / d->cp_d;
    const bool hyd = d->ref_state == CMDG_HYDROSTATIC_STATE && d->subtract_off;
    P.ref_sub = hyd;
    P.ref_src = hyd;
    P.gravity = d->gravity_source;
    P.turbulence = d->turbulence;
    P.with_div = d->with_divergence;
    P.rusanov = d->numerical_flux == CMDG_RUSANOV;
    P.naux = d->naux;
    P.a_Phi = d->orientation == CMDG_FLAT_ORIENTATION ? d->aux_Phi : -1;
    P.a_gPhi = d->orientation == CMDG_FLAT_ORIENTATION ? d->aux_gradPhi : -1;
    P.a_ref_rho = d->ref_state == CMDG_HYDROSTATIC_STATE ? d->aux_ref_rho : -1;
    P.a_ref_p = d->ref_state == CMDG_HYDROSTATIC_STATE ? d->aux_ref_p : -1;
    P.a_Delta = d->aux_Delta;
    P.a_theta_v = d->aux_theta_v;
    P.a_air_T = d->aux_air_T;

    const int Np = c->Np, Nfp = c->Nfp;
    const int64_t E = c->nelem;
    // connectivity: minus side must be the element's own face nodes (Grids.jl:615)
    const int Nq = c->Nq, NN = c->N;
    for (int64_t e = 0; e < (brick ? 0 : E); ++e)
        for (int f = 0; f < 6; ++f)
            for (int fn = 0; fn < Nfp; ++fn) {
                int a = fn % Nq, b = fn / Nq, vid;
                switch (f) {
                case 0: vid = 0 + Nq * (a + Nq * b); break;
                case 1: vid = NN + Nq * (a + Nq * b); break;
                case 2: vid = a + Nq * (0 + Nq * b); break;
                case 3: vid = a + Nq * (NN + Nq * b); break;
                case 4: vid = a + Nq * (b + Nq * 0); break;
                default: vid = a + Nq * (b + Nq * NN); break;
                }
                const int64_t id = fn + (int64_t)Nfp * (f + 6 * e);
                if (d->vmap_minus[id] - 1 != (int64_t)Np * e + vid)
                    return fail(c, CMDG_EINVAL, "vmap_minus is not the element's own face nodes");
                const int64_t p = d->vmap_plus[id] - 1;
                if (p < 0 || p >= (int64_t)Np * E) return fail(c, CMDG_EINVAL, "vmap_plus out of range");
            }
    int rc;
    if ((rc = upload(c, &c->d_D, d->D, (size_t)Nq * Nq))) return rc;
    if ((rc = upload(c, &c->d_aux, d->aux_init, (size_t)Np * (d->naux > 0 ? d->naux : 1) * E))) return rc;
    // the FP64 gradient flux of FP64 contexts (mixed ones keep theirs in FP32, d_gf32)
    if (d->precision != CMDG_MIXED && (rc = upload<double>(c, &c->d_gf, nullptr, (size_t)Np * c->ngf * E)))
        return rc;
    if ((rc = upload<double>(c, &c->d_Qs, nullptr, (size_t)Np * 5 * E))) return rc;
    if (c->d_gf &&
        (rc = check_hip(c, hipMemset(c->d_gf, 0, (size_t)Np * c->ngf * E * sizeof(double)), "hipMemset")))
        return rc;
    c->A.D = c->d_D;
    c->A.nelem = E;
    std::vector<int64_t> crep;  // first element of each static-field class (empty: no classes)
    {
        std::vector<double> geo, W, Wf;
        std::vector<int32_t> nbe;
        const char *env = std::getenv("CMDG_GENERAL_GEOMETRY");
        const bool force_general = env && env[0] == '1' && !brick;  // a brick has no per-node arrays
        std::vector<int64_t> lat[3];
        if (brick) {
            const std::string why = brick_geometry(d, brick, Nq, geo, nbe, W, Wf, lat);
            if (!why.empty()) return fail(c, CMDG_EINVAL, why);
        }
        if (brick || (!force_general && detect_affine(d, Nq, geo, nbe, W, Wf))) {
            c->affine = true;
            if ((rc = upload(c, &c->d_geo, geo.data(), geo.size()))) return rc;
            if ((rc = upload(c, &c->d_nbe, nbe.data(), nbe.size()))) return rc;
            if ((rc = upload(c, &c->d_W, W.data(), W.size()))) return rc;
            if ((rc = upload(c, &c->d_Wf, Wf.data(), Wf.size()))) return rc;
            // static face trace: Phi and p_ref at every face node
            std::vector<double> st((size_t)E * 6 * TR_S * Nfp, 0.0);
            for (int64_t e = 0; e < E; ++e)
                for (int f = 0; f < 6; ++f)
                    for (int fn = 0; fn < Nfp; ++fn) {
                        const int vid = fmask_vid(Nq, f, fn);
                        const double *ab = d->aux_init + (int64_t)Np * d->naux * e + vid;
                        const size_t o = (((size_t)e * 6 + f) * TR_S) * Nfp + fn;
                        st[o] = P.a_Phi >= 0 ? ab[(int64_t)Np * P.a_Phi] : -0.0;
                        st[o + Nfp] = P.a_ref_p >= 0 ? ab[(int64_t)Np * P.a_ref_p] : 0.0;
                    }
            // Phi and p_ref are functions of position in every model here; when they agree
            // bitwise on both sides of every interior face, the face kernels take the
            // neighbour's values from their own side and skip its static trace.
            bool cont = true;
            for (int64_t e = 0; e < E && cont; ++e)
                for (int f = 0; f < 6 && cont; ++f) {
                    const int32_t ne = nbe[6 * e + f];
                    if (ne < 0) continue;
                    const double *a = st.data() + (((size_t)e * 6 + f) * TR_S) * Nfp;
                    const double *b = st.data() + (((size_t)ne * 6 + (f ^ 1)) * TR_S) * Nfp;
                    cont = std::memcmp(a, b, sizeof(double) * TR_S * Nfp) == 0;
                }
            const char *tenv = std::getenv("CMDG_NEIGHBOUR_STATIC_TRACE");  // 1: always read it
            c->AA.aux_cont = (cont && !(tenv && tenv[0] == '1')) ? 1 : 0;
            if ((rc = upload(c, &c->d_str, st.data(), st.size()))) return rc;
            // the FP64 second-order face trace (mixed contexts carry the stress form instead)
            if (d->precision != CMDG_MIXED && c->diffusive &&
                (rc = upload<double>(c, &c->d_f2tr, nullptr, (size_t)E * 6 * TR_F2 * Nfp)))
                return rc;
            c->AA.geo = c->d_geo;
            c->AA.nbe = c->d_nbe;
            c->AA.W = c->d_W;
            c->AA.Wf = c->d_Wf;
            c->AA.D = c->d_D;
            c->AA.str = c->d_str;
            // static fields by class: on a brick with a reference state that depends on
            // height alone the kernels read one layer's fields, held in cache, instead of
            // 56 bytes a node (CMDG_STATIC_AUX_CLASSES=0 keeps one block per element)
            c->AA.sa = c->d_aux;
            c->AA.cls = nullptr;
            const char *cenv = std::getenv("CMDG_STATIC_AUX_CLASSES");
            std::vector<int32_t> cls;
            if (!(cenv && cenv[0] == '0') && d->naux > 0 && static_classes(d, P, Np, cls, crep)) {
                // the class's block of the aux layout; fields outside the class key are NaN
                const int64_t nc = (int64_t)crep.size(), na = d->naux;
                std::vector<double> sa((size_t)Np * na * nc, std::nan(""));
                for (int64_t k = 0; k < nc; ++k)
                    for (int f : {P.a_Phi, P.a_gPhi, P.a_gPhi >= 0 ? P.a_gPhi + 1 : -1,
                                  P.a_gPhi >= 0 ? P.a_gPhi + 2 : -1, P.a_ref_rho, P.a_ref_p, P.a_Delta})
                        if (f >= 0)
                            std::memcpy(&sa[(size_t)Np * (na * k + f)], d->aux_init + (int64_t)Np * (na * crep[k] + f),
                                        sizeof(double) * Np);
                if ((rc = upload(c, &c->d_cls, cls.data(), cls.size()))) return rc;
                if ((rc = upload(c, &c->d_sa, sa.data(), sa.size()))) return rc;
                c->AA.cls = c->d_cls;
                c->AA.sa = c->d_sa;
            } else {
                crep.clear();
            }
            // tiled traversal.  CMDG_ELEMENT_TILE=T sets cubic tiles of edge T (0 keeps
            // the storage order) and "TX,TY,TZ" tile extents along x, y, z (0: the whole
            // axis); CMDG_TILED_PASSES = tendency, gradient, both or none chooses the passes
            // that walk the tiles.  Default: the tendency pass walks 8^3 tiles.  A mixed
            // context whose face neighbours would, in storage order, leave one XCD's L2
            // before the gradient pass gathers their state (the 90th percentile of the
            // storage distance between face neighbours, in state bytes, above 4 MiB) walks
            // 12 x 12 columns of elements spanning the whole y axis, in both passes.
            // Measured on one box, chosen form (mixed + learned closure):
            //   96^3 (distance 9216 elements, 46 MB): tendency pass in 8^3 tiles 55.0 GDOF/s;
            //     both passes in 8^3 tiles 55.7-56.0 (edge 6 / 12 / 16: 55.3 / 56.2 / 55.8);
            //     both in 12 x 12 columns 56.6 (10 x 10: 56.1, 14 x 14: 55.8, 8 x 8: 56.1);
            //     mixed + Smagorinsky-Lilly 59.6 / 60.3 / 60.3;
            //   25^3 (625 elements, 3.1 MB): tendency pass in 8^3 tiles 50.0-50.3; both
            //     passes 49.3; 12 x 12 columns 49.2 (tendency pass), 48.6 (both).
            // FP64 at 96^3: 40.1-40.4 with the tendency pass in 8^3 tiles (12 x 12 columns
            // 40.2), 37.6-38.6 with both: its gradient pass, which writes the second-order
            // face trace, loses in tiles more than its gathers gain.
            bool far = false;
            if (d->precision == CMDG_MIXED) {
                std::vector<int64_t> gap;
                gap.reserve(nbe.size());
                for (int64_t e = 0; e < E; ++e)
                    for (int f = 0; f < 6; ++f)
                        if (nbe[6 * e + f] >= 0) gap.push_back(std::llabs((int64_t)nbe[6 * e + f] - e));
                if (!gap.empty()) {
                    auto q = gap.begin() + (int64_t)(0.9 * (double)(gap.size() - 1));
                    std::nth_element(gap.begin(), q, gap.end());
                    far = (double)*q * Np * 5 * sizeof(double) > 4.0 * 1024 * 1024;
                }
            }
            const char *tile_env = std::getenv("CMDG_ELEMENT_TILE");
            int T[3] = {8, 8, 8};
            if (far) T[0] = 12, T[1] = 0, T[2] = 12;
            if (tile_env && std::strchr(tile_env, ',')) {
                if (std::sscanf(tile_env, "%d,%d,%d", &T[0], &T[1], &T[2]) != 3) T[0] = T[1] = T[2] = -1;
            } else if (tile_env) {
                T[0] = T[1] = T[2] = std::atoi(tile_env) > 0 ? std::atoi(tile_env) : -1;
            }
            std::vector<int32_t> order;
            const int64_t nax[3] = {brick ? brick->ne[0] : 0, brick ? brick->ne[1] : 0, brick ? brick->ne[2] : 0};
            if (brick ? tiled_order_lattice(lat, nax, E, T, order) : tiled_order(d, Np, T, order)) {
                if ((rc = upload(c, &c->d_order, order.data(), order.size()))) return rc;
                const char *penv = std::getenv("CMDG_TILED_PASSES");
                const std::string passes = penv ? penv : (far ? "both" : "tendency");
                if (passes == "both" || passes == "gradient") c->ord_grad = c->d_order;
                if (passes == "both" || passes == "tendency") c->ord_tend = c->d_order;
            }
        }
    }
    if (!c->affine) {
        // general meshes: per-node geometry xi (9), M, MI, the face geometry and the
        // neighbour node of every face node (affine bricks read none of these)
        std::vector<double> vg((size_t)Np * NVG * E);
        for (int64_t e = 0; e < E; ++e)
            for (int f = 0; f < NVG; ++f)
                std::memcpy(&vg[(size_t)Np * (f + NVG * e)], d->vgeo + (size_t)Np * (f + (int64_t)d->nvgeo * e),
                            Np * sizeof(double));
        std::vector<int32_t> nbr((size_t)Nfp * 6 * E), bnd((size_t)6 * E);
        for (int64_t i = 0; i < (int64_t)Nfp * 6 * E; ++i) nbr[i] = (int32_t)(d->vmap_plus[i] - 1);
        for (int64_t i = 0; i < 6 * E; ++i) bnd[i] = (int32_t)d->elemtobndy[i];
        if ((rc = upload(c, &c->d_vg, vg.data(), vg.size()))) return rc;
        if ((rc = upload(c, &c->d_sg, d->sgeo, (size_t)5 * Nfp * 6 * E))) return rc;
        if ((rc = upload(c, &c->d_nbr, nbr.data(), nbr.size()))) return rc;
        if ((rc = upload(c, &c->d_bnd, bnd.data(), bnd.size()))) return rc;
        c->A.vg = c->d_vg;
        c->A.sg = c->d_sg;
        c->A.nbr = c->d_nbr;
        c->A.bnd = c->d_bnd;
    }
    if (d->precision == CMDG_MIXED) {
        if (!c->affine)
            return fail(c, CMDG_EUNSUPPORTED, "mixed precision runs on axis-aligned affine brick meshes only");
        c->mixed = true;
        MixParams &M = c->MP;
        M.R_d = (float)d->R_d;
        M.gammaR = (float)P.gammaR;
        M.kappa = (float)P.kappa;
        M.inv_MSLP = (float)(1.0 / d->MSLP);
        M.grav = (float)d->grav;
        M.C_smag = (float)d->C_smag;
        M.inv_Pr = (float)d->inv_Pr_turb;
        M.visc = (float)d->viscosity;
        M.dyn_visc = d->turbulence == CMDG_CONSTANT_DYNAMIC_VISCOSITY;
        M.with_div = d->with_divergence;
        M.ref_sub = P.ref_sub;
        M.ref_src = P.ref_src;
        M.gravity = P.gravity;
        M.rusanov = P.rusanov;
        c->iso32 = !(c->smag && P.a_gPhi >= 0);
        c->ngf32 = c->iso32 ? StressForm<true>::NGF : StressForm<false>::NGF;
        if ((rc = upload<float>(c, &c->d_gf32, nullptr, (size_t)Np * c->ngf32 * E))) return rc;
        if ((rc = check_hip(c, hipMemset(c->d_gf32, 0, (size_t)Np * c->ngf32 * E * sizeof(float)), "hipMemset")))
            return rc;
        if (P.a_gPhi >= 0) {
            // grad Phi enters the mixed kernels in FP32 only: keep it rounded, 12 B a node
            // (per static-field class when there are classes)
            const int64_t nb = crep.empty() ? E : (int64_t)crep.size();
            std::vector<float> g((size_t)Np * 3 * nb);
            for (int64_t b = 0; b < nb; ++b) {
                const int64_t e = crep.empty() ? b : crep[b];
                for (int a = 0; a < 3; ++a)
                    for (int n = 0; n < Np; ++n)
                        g[((size_t)b * 3 + a) * Np + n] =
                            (float)d->aux_init[(int64_t)Np * ((int64_t)d->naux * e + P.a_gPhi + a) + n];
            }
            if ((rc = upload(c, &c->d_gphi32, g.data(), g.size()))) return rc;
            c->AA.gphi32 = c->d_gphi32;
        }
        if (c->learned && (rc = upload<MlpDev>(c, &c->d_mlp, nullptr, 1))) return rc;
    }
    return CMDG_OK;
}

}  // namespace

extern "C" {

int cmdg_destroy(cmdg_ctx *c) {
    if (!c) return CMDG_OK;
    for (auto &r : c->recs) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    (void)hipFree(c->d_vg);
    (void)hipFree(c->d_sg);
    (void)hipFree(c->d_nbr);
    (void)hipFree(c->d_bnd);
    (void)hipFree(c->d_D);
    (void)hipFree(c->d_aux);
    (void)hipFree(c->d_gf);
    (void)hipFree(c->d_Qs);
    (void)hipFree(c->d_geo);
    (void)hipFree(c->d_nbe);
    (void)hipFree(c->d_W);
    (void)hipFree(c->d_Wf);
    (void)hipFree(c->d_f2tr);
    (void)hipFree(c->d_str);
    (void)hipFree(c->d_cls);
    (void)hipFree(c->d_sa);
    (void)hipFree(c->d_order);
    (void)hipFree(c->d_gf32);
    (void)hipFree(c->d_gphi32);
    (void)hipFree(c->d_mlp);
    (void)hipFree(c->d_part);
    (void)hipFree(c->d_scalar);
    delete c;
    return CMDG_OK;
}

int cmdg_rhs(cmdg_ctx *c, double *tend, const double *Q, double t, double alpha, double beta) {
    (void)t;
    if (!c || !tend || !Q) return c ? fail(c, CMDG_EINVAL, "null pointer") : CMDG_EINVAL;
    if (tend == Q) return fail(c, CMDG_EINVAL, "tendency aliases the state");
    return launch_rhs(c, Q, tend, alpha, beta);
}

int cmdg_lsrk_update(double *dQ, double *Q, int64_t n, double rka, double rkb, double dt, void *stream) {
    if (!dQ || !Q || n < 0) return CMDG_EINVAL;
    if (n == 0) return CMDG_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_lsrk_update, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, dQ, Q, n, rka, rkb * dt);
    return hipGetLastError() == hipSuccess ? CMDG_OK : CMDG_EHIP;
}

int cmdg_lsrk_stage(cmdg_ctx *c, double *Q, double *dQ, double t_stage, double dt, double rka_next, double rkb) {
    (void)t_stage;
    if (!c || !Q || !dQ) return c ? fail(c, CMDG_EINVAL, "null pointer") : CMDG_EINVAL;
    int rc = launch_rhs(c, Q, dQ, 1.0, 1.0);
    if (rc) return rc;
    {
        TimedLaunch tl(c, KIND_UPDATE);
        rc = cmdg_lsrk_update(dQ, Q, (int64_t)c->Np * 5 * c->nelem, rka_next, rkb, dt, c->stream);
    }
    return rc ? fail(c, rc, "update launch") : CMDG_OK;
}

int cmdg_lsrk54_steps(cmdg_ctx *c, double *Q, double *dQ, double t, double dt, int64_t nsteps) {
    (void)t;
    if (!c || !Q || !dQ) return c ? fail(c, CMDG_EINVAL, "null pointer") : CMDG_EINVAL;
    if (nsteps < 0) return fail(c, CMDG_EINVAL, "negative step count");
    if (c->learned && !c->mlp_ready) return fail(c, CMDG_EINVAL, "learned closure weights not set");
    if (nsteps == 0) return CMDG_OK;
    int rc;
    // Every stage reads its neighbours' volume state, so it writes the next stage's
    // state to another buffer: the stages alternate between the caller's Q (C) and the
    // context buffer S, and no caller code runs between the steps of one call.  Over
    // 5 nsteps stages the result lands in C when that count is even; when it is odd,
    // S <- C first (on general meshes by the first gradient pass) and the first stage
    // reads S.
    double *C = Q, *S = c->d_Qs;
    double *src = C, *dst = S;
    const bool odd = (nsteps % 2) != 0;
    if (odd) {
        src = S;
        dst = C;
        if ((c->affine || !c->diffusive) && (rc = launch_copy(c, S, C, (int64_t)c->Np * 5 * c->nelem))) return rc;
    }
    for (int64_t step = 0; step < nsteps; ++step)
        for (int s = 0; s < 5; ++s) {
            const double a = RKA[(s + 1) % 5], bdt = RKB[s] * dt;
            // aux keeps theta_v and air_T of the last stage's state, as after dostep!.  No
            // kernel reads them back (each forms them from the state), so only the last
            // step's write reaches the caller and the earlier steps skip it (theta_v costs
            // a power per node).
            const int wa = (s == 4 && step == nsteps - 1);
            // The last stage scales dQ by RKA[0] = 0 and the next step's first stage adds
            // the right-hand side to those zeros: between two steps of the call the zeros
            // are neither stored nor read (dQ is read on entry and written on return as
            // by nsteps dostep! calls; a zero in the state may differ in sign).
            const double dq_read = (s == 0 && step > 0) ? 0.0 : 1.0;
            const int dq_store = !(s == 4 && step < nsteps - 1);
            if (c->diffusive) {
                const bool first = odd && step == 0 && s == 0 && !c->affine;
                if ((rc = launch_gradient(c, first ? C : src, first ? S : nullptr, wa))) return rc;
            }
            if ((rc = launch_tendency(c, src, dQ, 1.0, dq_read, 1, dst, a, bdt, c->diffusive ? 0 : wa, dq_store)))
                return rc;
            double *tmp = src;
            src = dst;
            dst = tmp;
        }
    return CMDG_OK;
}

int cmdg_lsrk54_step(cmdg_ctx *c, double *Q, double *dQ, double t, double dt) {
    return cmdg_lsrk54_steps(c, Q, dQ, t, dt, 1);
}

int cmdg_aux(cmdg_ctx *c, double **aux) {
    if (!c || !aux) return CMDG_EINVAL;
    *aux = c->d_aux;
    return CMDG_OK;
}

int cmdg_gradient_flux(cmdg_ctx *c, double **gf, int32_t *ngf) {
    if (!c || !gf) return CMDG_EINVAL;
    if (c->mixed) return fail(c, CMDG_EUNSUPPORTED, "the mixed-precision gradient flux is stored in FP32");
    *gf = c->d_gf;
    if (ngf) *ngf = c->ngf;
    return CMDG_OK;
}

int cmdg_gradient_flux_f32(cmdg_ctx *c, float **gf, int32_t *ngf) {
    if (!c || !gf) return CMDG_EINVAL;
    if (!c->mixed) return fail(c, CMDG_EUNSUPPORTED, "the FP64 path stores its gradient flux in FP64");
    *gf = c->d_gf32;
    if (ngf) *ngf = c->ngf32;
    return CMDG_OK;
}

int cmdg_mlp_set_weights(cmdg_ctx *c, const uint16_t *w1, const float *b1, const uint16_t *w2, const float *b2,
                         const uint16_t *w3, const float *b3, const float *shift, const float *scale) {
    if (!c) return CMDG_EINVAL;
    if (!c->learned) return fail(c, CMDG_EINVAL, "the context was not created with the learned closure");
    if (!w1 || !b1 || !w2 || !b2 || !w3 || !b3 || !shift || !scale) return fail(c, CMDG_EINVAL, "null pointer");
    const MlpDev h = mlp_device_copy(w1, b1, w2, b2, w3, b3, shift, scale);
    int rc = check_hip(c, hipMemcpyAsync(c->d_mlp, &h, sizeof(h), hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync");
    if (rc) return rc;
    rc = check_hip(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (!rc) c->mlp_ready = true;
    return rc;
}

int cmdg_mlp_forward(cmdg_ctx *c, const float *x_raw, float *y, int64_t n) {
    if (!c) return CMDG_EINVAL;
    if (!c->learned || !c->mlp_ready) return fail(c, CMDG_EINVAL, "learned closure weights not set");
    if (n < 0 || (n > 0 && (!x_raw || !y))) return fail(c, CMDG_EINVAL, "bad arguments");
    if (n == 0) return CMDG_OK;
    const int64_t nb = (n + 127) / 128;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)nb), dim3(128), 0, c->stream, c->d_mlp, x_raw, y, n);
    return check_hip(c, hipGetLastError(), "mlp forward launch");
}

int cmdg_velocity_gradient(cmdg_ctx *c, const double *Q, double *grad_u) {
    if (!c) return CMDG_EINVAL;
    if (!Q || !grad_u) return fail(c, CMDG_EINVAL, "null pointer");
    if (!c->affine) return fail(c, CMDG_EUNSUPPORTED, "the velocity gradient is built for affine brick meshes");
    dim3 grid((unsigned)xcd_grid(c->nelem));
    if (c->Nq == 5)
        hipLaunchKernelGGL(k_velgrad_aff<5>, grid, dim3(threads_for<5>()), 0, c->stream, c->AA, c->nelem, Q, grad_u);
    else
        hipLaunchKernelGGL(k_velgrad_aff<4>, grid, dim3(threads_for<4>()), 0, c->stream, c->AA, c->nelem, Q, grad_u);
    return check_hip(c, hipGetLastError(), "velocity gradient launch");
}

int cmdg_box_filter(const double *Qf, const int32_t *children, int64_t ncoarse, int32_t N, const double *proj,
                    double *moments, void *stream) {
    if ((N != 3 && N != 4) || ncoarse < 0 || !proj) return CMDG_EINVAL;
    if (ncoarse == 0) return CMDG_OK;
    if (!Qf || !children || !moments) return CMDG_EINVAL;
    const int Nq = N + 1;
    FilterProj PR{};
    for (int h = 0; h < 2; ++h)
        for (int i = 0; i < Nq; ++i)
            for (int a = 0; a < Nq; ++a) PR.p[h][i][a] = proj[(h * Nq + i) * Nq + a];
    if (Nq == 5)
        hipLaunchKernelGGL(k_box_filter<5>, dim3((unsigned)ncoarse), dim3(threads_for<5>()), 0, (hipStream_t)stream, Qf,
                           children, ncoarse, PR, moments);
    else
        hipLaunchKernelGGL(k_box_filter<4>, dim3((unsigned)ncoarse), dim3(threads_for<4>()), 0, (hipStream_t)stream, Qf,
                           children, ncoarse, PR, moments);
    return hipGetLastError() == hipSuccess ? CMDG_OK : CMDG_EHIP;
}

int cmdg_closure_samples(const double *moments, const double *grad_u, int64_t nelem, int32_t N, double Delta,
                         double nu_mol, double nu_floor, int32_t element_average, float *samples, int64_t ld,
                         int64_t offset, void *stream) {
    if ((N != 3 && N != 4) || nelem < 0 || offset < 0) return CMDG_EINVAL;
    const int Np = (N + 1) * (N + 1) * (N + 1);
    const int64_t n = nelem * Np;
    if (offset + n > ld) return CMDG_EINVAL;
    if (n == 0) return CMDG_OK;
    if (!moments || !grad_u || !samples) return CMDG_EINVAL;
    hipLaunchKernelGGL(k_closure_samples, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, moments,
                       grad_u, nelem, Np, Delta, nu_mol, nu_floor, samples, ld, offset);
    if (hipGetLastError() != hipSuccess) return CMDG_EHIP;
    if (element_average)
        hipLaunchKernelGGL(k_element_average_pqr, dim3((unsigned)nelem), dim3(128), 0, (hipStream_t)stream, samples, ld,
                           offset, Np);
    return hipGetLastError() == hipSuccess ? CMDG_OK : CMDG_EHIP;
}

int cmdg_mlp_loss_grad(const float *params, const float *shift, const float *scale, const float *samples, int64_t ld,
                       const int64_t *idx, int64_t batch, double inv_norm, float *grad, float *work,
                       int64_t work_floats, void *stream) {
    if (batch <= 0 || !params || !shift || !scale || !samples || !grad || !work) return CMDG_EINVAL;
    if (!idx && batch > ld) return CMDG_EINVAL;
    const int64_t nb = (batch + TRAIN_TB - 1) / TRAIN_TB;
    if (nb * GRAD_W > work_floats) return CMDG_EINVAL;
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)nb), dim3(TRAIN_TB), 0, (hipStream_t)stream, params, shift, scale,
                       samples, ld, idx, batch, (float)inv_norm, work);
    if (hipGetLastError() != hipSuccess) return CMDG_EHIP;
    hipLaunchKernelGGL(k_grad_reduce, dim3((GRAD_W + 255) / 256), dim3(256), 0, (hipStream_t)stream, work, nb, grad);
    return hipGetLastError() == hipSuccess ? CMDG_OK : CMDG_EHIP;
}

}  // extern "C"

// Partial gradients of cmdg_mlp_loss_grad_static: one GRAD_W row per workgroup of
// TRAIN_TB samples, for at most MLP_STATIC_BLOCKS workgroups.  Each GPU holds its own
// copy; k_mlp_grad writes every entry of the rows k_grad_reduce reads, so nothing in
// it has to start at zero.
constexpr int64_t MLP_STATIC_BLOCKS = CMDG_MLP_STATIC_MAX_BATCH / TRAIN_TB;
static_assert(MLP_STATIC_BLOCKS * TRAIN_TB == CMDG_MLP_STATIC_MAX_BATCH, "whole workgroups");
static __device__ __align__(256) float scratch[MLP_STATIC_BLOCKS * GRAD_W];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on the device of `stream` (the current device for
// the null stream), looked up once per device with that device current.
int static_scratch(hipStream_t stream, float **out) {
    int dev = -1, cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return CMDG_EHIP;
    dev = cur;
    if (stream && hipStreamGetDevice(stream, &dev) != hipSuccess) return CMDG_EHIP;
    if (dev < 0 || dev >= kMaxDevices) return CMDG_EUNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        if (dev != cur && hipSetDevice(dev) != hipSuccess) return CMDG_EHIP;
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (dev != cur && hipSetDevice(cur) != hipSuccess) return CMDG_EHIP;
        if (e != hipSuccess || !p) return CMDG_EHIP;
        g_scratch_dev[dev] = static_cast<float *>(p);
    }
    *out = g_scratch_dev[dev];
    return CMDG_OK;
}

}  // namespace

extern "C" {

int cmdg_mlp_loss_grad_static(const float *params, const float *shift, const float *scale, const float *samples,
                              int64_t ld, const int64_t *idx, int64_t batch, double inv_norm, float *grad,
                              void *stream) {
    if (batch <= 0 || !params || !shift || !scale || !samples || !grad) return CMDG_EINVAL;
    if (!idx && batch > ld) return CMDG_EINVAL;
    if (batch > CMDG_MLP_STATIC_MAX_BATCH) return CMDG_ERANGE;
    float *work = nullptr;
    const int rc = static_scratch((hipStream_t)stream, &work);
    if (rc) return rc;
    const int64_t nb = (batch + TRAIN_TB - 1) / TRAIN_TB;
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)nb), dim3(TRAIN_TB), 0, (hipStream_t)stream, params, shift, scale,
                       samples, ld, idx, batch, (float)inv_norm, work);
    if (hipGetLastError() != hipSuccess) return CMDG_EHIP;
    hipLaunchKernelGGL(k_grad_reduce, dim3((GRAD_W + 255) / 256), dim3(256), 0, (hipStream_t)stream, work, nb, grad);
    return hipGetLastError() == hipSuccess ? CMDG_OK : CMDG_EHIP;
}

int cmdg_mlp_adam(float *params, const float *grad, float *m, float *v, int64_t n, float lr, float beta1, float beta2,
                  float eps, int64_t step, void *stream) {
    if (n <= 0 || step < 1 || !params || !grad || !m || !v) return CMDG_EINVAL;
    const float c1 = (float)(1.0 - std::pow((double)beta1, (double)step));
    const float c2 = (float)(1.0 - std::pow((double)beta2, (double)step));
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, params, grad, m, v, n,
                       lr, beta1, beta2, eps, c1, c2);
    return hipGetLastError() == hipSuccess ? CMDG_OK : CMDG_EHIP;
}

int cmdg_interpolate_equispaced(const double *Q, const int32_t *eijk, int64_t nelem, int32_t N, int32_t p,
                                const double *L, int64_t nx, int64_t ny, int64_t nz, double *u, void *stream) {
    if ((N != 3 && N != 4) || p < 1 || p > 8 || nelem < 0 || !L) return CMDG_EINVAL;
    if (nelem == 0) return CMDG_OK;
    if (!Q || !eijk || !u || nx <= 0 || ny <= 0 || nz <= 0) return CMDG_EINVAL;
    const int Nq = N + 1;
    InterpMat IM{};
    for (int m = 0; m < p; ++m)
        for (int i = 0; i < Nq; ++i) IM.L[m][i] = L[m * Nq + i];
    const unsigned nt = (unsigned)(((p * p * p) + 63) / 64 * 64);
    if (Nq == 5)
        hipLaunchKernelGGL(k_interp_equispaced<5>, dim3((unsigned)nelem), dim3(nt < 128 ? 128 : nt), 0,
                           (hipStream_t)stream, Q, eijk, nelem, p, IM, nx, ny, nz, u);
    else
        hipLaunchKernelGGL(k_interp_equispaced<4>, dim3((unsigned)nelem), dim3(nt < 64 ? 64 : nt), 0,
                           (hipStream_t)stream, Q, eijk, nelem, p, IM, nx, ny, nz, u);
    return hipGetLastError() == hipSuccess ? CMDG_OK : CMDG_EHIP;
}

static int diag_scratch(cmdg_ctx *c, int64_t n) {
    if (!c->d_scalar) {
        int rc = check_hip(c, hipMalloc((void **)&c->d_scalar, sizeof(double)), "hipMalloc");
        if (rc) return rc;
    }
    if (c->npart < n) {
        (void)hipFree(c->d_part);
        c->d_part = nullptr;
        c->npart = 0;
        int rc = check_hip(c, hipMalloc((void **)&c->d_part, (size_t)n * sizeof(double2)), "hipMalloc");
        if (rc) return rc;
        c->npart = n;
    }
    return CMDG_OK;
}

static int diag_result(cmdg_ctx *c, double *out) {
    int rc = check_hip(c, hipGetLastError(), "diagnostic launch");
    if (rc) return rc;
    rc = check_hip(c, hipMemcpyAsync(out, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync");
    if (rc) return rc;
    return check_hip(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

int cmdg_courant(cmdg_ctx *c, const double *Q, double dt, int32_t kind, int32_t direction, double *out) {
    if (!c) return CMDG_EINVAL;
    if (!Q || !out) return fail(c, CMDG_EINVAL, "null pointer");
    if (kind < 0 || kind > 2 || direction < 0 || direction > 2) return fail(c, CMDG_EINVAL, "bad Courant kind or direction");
    if (c->aux_coord < 0) return fail(c, CMDG_EINVAL, "the Courant number needs coordinates in aux");
    int rc = diag_scratch(c, c->nelem);
    if (rc) return rc;
    double *part = reinterpret_cast<double *>(c->d_part);
    const bool diff = kind == COURANT_DIFFUSIVE && c->diffusive;
    const double *gf = (diff && !c->mixed) ? c->d_gf : nullptr;
    const float *gf32 = (diff && c->mixed) ? c->d_gf32 : nullptr;
    const int ngf = diff ? (c->mixed ? c->ngf32 : c->ngf) : 0;
    dim3 grid((unsigned)c->nelem);
#define CMDG_COURANT(NQ, SM, MIX, ISO)                                                                          \
    hipLaunchKernelGGL((k_courant<NQ, SM, MIX, ISO>), grid, dim3(threads_for<NQ>()), 0, c->stream, c->P, c->MP, \
                       Q, c->d_aux, gf, gf32, ngf, c->nelem, dt, kind, direction, c->aux_coord, part)
#define CMDG_COURANT_NQ(NQ)                                             \
    do {                                                                \
        if (gf32 && c->smag && c->iso32) CMDG_COURANT(NQ, true, true, true); \
        else if (gf32 && c->smag) CMDG_COURANT(NQ, true, true, false);  \
        else if (gf32) CMDG_COURANT(NQ, false, true, true);             \
        else if (c->smag) CMDG_COURANT(NQ, true, false, true);          \
        else CMDG_COURANT(NQ, false, false, true);                      \
    } while (0)
    if (c->Nq == 5) CMDG_COURANT_NQ(5);
    else CMDG_COURANT_NQ(4);
#undef CMDG_COURANT_NQ
#undef CMDG_COURANT
    hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(256), 0, c->stream, part, c->nelem, c->d_scalar);
    return diag_result(c, out);
}

int cmdg_weightedsum(cmdg_ctx *c, const double *Q, int32_t state, double *out) {
    if (!c) return CMDG_EINVAL;
    if (!Q || !out) return fail(c, CMDG_EINVAL, "null pointer");
    if (state < 0 || state > 4) return fail(c, CMDG_EINVAL, "state index out of range");
    const int64_t total = (int64_t)c->Np * c->nelem;
    int64_t nb = (total + 255) / 256;
    if (nb > 1024) nb = 1024;
    int rc = diag_scratch(c, nb);
    if (rc) return rc;
    // M per node: vgeo's mass on general meshes; J w_i w_j w_k on affine bricks, which
    // the create-time check found equal to vgeo's M bit for bit
    hipLaunchKernelGGL(k_weightedsum, dim3((unsigned)nb), dim3(256), 0, c->stream, Q, c->d_vg, c->Np, NVG,
                       c->affine ? c->d_geo : nullptr, c->d_W, state, c->nelem, c->d_part);
    hipLaunchKernelGGL(k_dd_reduce, dim3(1), dim3(256), 0, c->stream, c->d_part, (int)nb, c->d_scalar);
    return diag_result(c, out);
}

int cmdg_not_finite_fields(cmdg_ctx *c, const double *Q, int32_t *mask) {
    if (!c) return CMDG_EINVAL;
    if (!Q || !mask) return fail(c, CMDG_EINVAL, "null pointer");
    const int64_t total = (int64_t)c->Np * 5 * c->nelem;
    int64_t nb = (total + 255) / 256;
    if (nb > 1024) nb = 1024;
    int rc = diag_scratch(c, nb);
    if (rc) return rc;
    hipLaunchKernelGGL(k_not_finite, dim3((unsigned)nb), dim3(256), 0, c->stream, Q, c->Np, 5, total, c->d_part);
    hipLaunchKernelGGL(k_or_reduce, dim3(1), dim3(256), 0, c->stream, c->d_part, (int)nb, c->d_scalar);
    double m = 0.0;
    rc = diag_result(c, &m);
    if (rc) return rc;
    *mask = (int32_t)m;
    return CMDG_OK;
}

int cmdg_geometry_kind(cmdg_ctx *c, int32_t *kind) {
    if (!c || !kind) return CMDG_EINVAL;
    *kind = (c->affine ? 1 : 0) | (c->affine && c->AA.aux_cont ? 2 : 0) | (c->affine && c->AA.cls ? 4 : 0) |
            (c->affine && (c->ord_grad || c->ord_tend) ? 8 : 0);
    return CMDG_OK;
}

int cmdg_set_kernel_timing(cmdg_ctx *c, int enable) {
    if (!c) return CMDG_EINVAL;
    c->timing = enable != 0;
    return CMDG_OK;
}

int cmdg_kernel_timing(cmdg_ctx *c, double *ms_total, int64_t *launches, int32_t nkinds) {
    if (!c || !ms_total || !launches || nkinds < 1) return CMDG_EINVAL;
    for (int k = 0; k < nkinds; ++k) { ms_total[k] = 0.0; launches[k] = 0; }
    int rc = check_hip(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (rc) return rc;
    for (auto &r : c->recs) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.start, r.stop) == hipSuccess && r.kind < nkinds) {
            ms_total[r.kind] += ms;
            launches[r.kind] += 1;
        }
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    c->recs.clear();
    return CMDG_OK;
}

int cmdg_sync(cmdg_ctx *c) {
    if (!c) return CMDG_EINVAL;
    return check_hip(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

}  // extern "C"
