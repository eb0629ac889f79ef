// count_flops.cpp -- TEST INFRASTRUCTURE ONLY.  The oracle (cmdg_oracle.cpp, included
// below) instantiated with an operation-counting scalar: every add, subtract, multiply,
// divide, square root and transcendental that the restated reference algorithm performs
// on its floating-point values is counted.  One right-hand side on a small periodic mesh
// divided by its node count gives the algorithmic operation count per node and stage,
// split at the two passes the HIP path launches: the gradient pass (auxiliary refresh,
// volume and interface gradients, learned nu_0) and the tendency pass (volume and
// interface tendency, LSRK update).  Built without OpenMP: the counters are global.
#include <cstdint>
#include <cmath>
#include <limits>

namespace opcount {
struct Ops {
    uint64_t add = 0, mul = 0, div = 0, sqrt = 0, trans = 0, cmp = 0;
};
inline Ops g;
}  // namespace opcount

struct Cnt {
    double v;
    Cnt() : v(0.0) {}
    Cnt(double x) : v(x) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    Cnt operator-() const { return Cnt(-v); }
    Cnt &operator+=(Cnt o) { ++opcount::g.add; v += o.v; return *this; }
    Cnt &operator-=(Cnt o) { ++opcount::g.add; v -= o.v; return *this; }
    Cnt &operator*=(Cnt o) { ++opcount::g.mul; v *= o.v; return *this; }
    Cnt &operator/=(Cnt o) { ++opcount::g.div; v /= o.v; return *this; }
};
inline Cnt operator+(Cnt a, Cnt b) { ++opcount::g.add; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { ++opcount::g.add; return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { ++opcount::g.mul; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { ++opcount::g.div; return Cnt(a.v / b.v); }
inline bool operator<(Cnt a, Cnt b) { ++opcount::g.cmp; return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { ++opcount::g.cmp; return a.v > b.v; }
inline bool operator<=(Cnt a, Cnt b) { ++opcount::g.cmp; return a.v <= b.v; }
inline bool operator>=(Cnt a, Cnt b) { ++opcount::g.cmp; return a.v >= b.v; }
inline bool operator==(Cnt a, Cnt b) { ++opcount::g.cmp; return a.v == b.v; }
inline bool operator!=(Cnt a, Cnt b) { ++opcount::g.cmp; return a.v != b.v; }

namespace std {
inline Cnt fabs(Cnt a) { return Cnt(std::fabs(a.v)); }
inline Cnt sqrt(Cnt a) { ++opcount::g.sqrt; return Cnt(std::sqrt(a.v)); }
inline Cnt pow(Cnt a, Cnt b) { ++opcount::g.trans; return Cnt(std::pow(a.v, b.v)); }
inline Cnt exp(Cnt a) { ++opcount::g.trans; return Cnt(std::exp(a.v)); }
inline Cnt nextafter(Cnt a, Cnt b) { return Cnt(std::nextafter(a.v, b.v)); }
template <> struct numeric_limits<Cnt> {
    static Cnt infinity() { return Cnt(std::numeric_limits<double>::infinity()); }
};
}  // namespace std

#include "cmdg_oracle.cpp"

namespace {
void snap(double *o) {
    o[0] = (double)opcount::g.add;
    o[1] = (double)opcount::g.mul;
    o[2] = (double)opcount::g.div;
    o[3] = (double)opcount::g.sqrt;
    o[4] = (double)opcount::g.trans;
    o[5] = (double)opcount::g.cmp;
    opcount::g = opcount::Ops{};
}
}  // namespace

extern "C" {
// Operation counts of one right-hand side plus one LSRK update of the state Q with
// auxiliary state aux (both (Np, *, nelem), FP64): out[0..5] the gradient pass,
// out[6..11] the tendency pass, each {add/sub, mul, div, sqrt, pow/exp, compare}.
int oracle_count_ops(const cmdg_desc *d, const double *Q, const double *aux_in, double *out) {
    Ctx<Cnt> *c = make_ctx<Cnt>(d);
    const int64_t nq = (int64_t)c->Np * 5 * c->nelem, na = (int64_t)c->Np * c->naux * c->nelem;
    std::vector<Cnt> q(Q, Q + nq), aux(aux_in, aux_in + na), dq(nq, Cnt(0.0));
    const Cnt a = rkcoef<Cnt>(0, 1), bdt = rkcoef<Cnt>(1, 0) * Cnt(1e-3);  // per stage, not per node
    opcount::g = opcount::Ops{};
    update_aux(*c, q.data(), aux.data());
    if (c->diffusive) {
        volume_gradients(*c, q.data(), aux.data());
        interface_gradients(*c, q.data(), aux.data());
        if (c->learned) learned_viscosity(*c, aux.data());
    }
    snap(out);
    volume_tendency(*c, dq.data(), q.data(), aux.data(), Cnt(1.0), Cnt(1.0));
    interface_tendency(*c, dq.data(), q.data(), aux.data(), Cnt(1.0));
    for (int64_t i = 0; i < nq; ++i) {
        q[i] += bdt * dq[i];  // update! (LowStorageRungeKuttaMethod.jl:146-158)
        dq[i] *= a;
    }
    snap(out + 6);
    delete c;
    return 0;
}
}  // extern "C"
