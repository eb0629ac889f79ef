// cmdg_general.h -- DG kernels for general (curvilinear) hexahedral meshes: per-node
// metric terms, per-face-node normals and surface weights, vmap+ gathers.
#pragma once
#include "cmdg_physics.h"

namespace cmdg {

// ------------------------------------------------------------------ gradient kernel
template <int Nq, bool SMAG>
__global__ void __launch_bounds__(threads_for<Nq>())
k_gradient(KParams P, KArrays A, const double *__restrict__ Q, double *__restrict__ aux,
           double *__restrict__ gfo, double *__restrict__ Qcopy, int write_aux) {
    constexpr int Np = Nq * Nq * Nq, Nfp = Nq * Nq;
    constexpr int NG = SMAG ? 5 : 4, NGF = SMAG ? 10 : 9;
    __shared__ double sD[Nq * Nq];
    __shared__ double sG[NG][Np];
    const int64_t e = xcd_element(blockIdx.x, A.nelem);
    if (e >= A.nelem) return;
    const int n = threadIdx.x;
    if (n < Nq * Nq) sD[n] = A.D[n];
    const bool act = n < Np;
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
    double q[5], thv = 0.0;
    Aux a;
    Thermo t;
    if (act) {
        const double *qb = Q + (int64_t)Np * 5 * e + n;
#pragma unroll
        for (int s = 0; s < 5; ++s) q[s] = qb[Np * s];
        if (Qcopy) {
            double *qc = Qcopy + (int64_t)Np * 5 * e + n;
#pragma unroll
            for (int s = 0; s < 5; ++s) qc[Np * s] = q[s];
        }
        a = load_aux(P, aux, Np, n, e);
        t = thermo(P, q, a.Phi);
        thv = theta_v(P, t);
        if (write_aux) {
            double *ab = aux + (int64_t)Np * P.naux * e + n;
            if (P.a_theta_v >= 0) ab[(int64_t)Np * P.a_theta_v] = thv;
            if (P.a_air_T >= 0) ab[(int64_t)Np * P.a_air_T] = t.temp;
        }
        double G[5];
        grad_arg<SMAG>(P, q, t, thv, G);
#pragma unroll
        for (int s = 0; s < NG; ++s) sG[s][n] = G[s];
    }
    __syncthreads();
    if (!act) return;

    const double *vg = A.vg + (int64_t)Np * NVG * e + n;
    double xi[9];
#pragma unroll
    for (int f = 0; f < 9; ++f) xi[f] = vg[Np * f];

    // horizontal pass, then vertical pass (DGModel_kernels.jl:1007-1127, 1235-1325):
    // the gradient-flux transform is applied to each pass and the results summed
    double gG[3][5], gV[3][5];
#pragma unroll
    for (int s = 0; s < 5; ++s)
#pragma unroll
        for (int d = 0; d < 3; ++d) gG[d][s] = gV[d][s] = 0.0;
#pragma unroll
    for (int s = 0; s < NG; ++s) {
        double g1 = 0.0, g2 = 0.0, g3 = -0.0;
#pragma unroll
        for (int m = 0; m < Nq; ++m) {
            g1 += sD[i + Nq * m] * sG[s][m + Nq * (j + Nq * k)];
            g2 += sD[j + Nq * m] * sG[s][i + Nq * (m + Nq * k)];
        }
#pragma unroll
        for (int m = 0; m < Nq; ++m) g3 += sD[k + Nq * m] * sG[s][i + Nq * (j + Nq * m)];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double v = -0.0;
            v += xi[0 + 3 * d] * g1;
            v += xi[1 + 3 * d] * g2;
            gG[d][s] = v;
            gV[d][s] = -0.0 + xi[2 + 3 * d] * g3;
        }
    }

    double gf[10], gfv[10];
    grad_flux<SMAG>(gG, a, thv, gf);
    grad_flux<SMAG>(gV, a, thv, gfv);
#pragma unroll
    for (int s = 0; s < NGF; ++s) gf[s] += gfv[s];

    // lifted face jumps: M^-1 Mf (Phi(G*) - Phi(n G-)), CentralNumericalFluxGradient
    double Gm[5];
    grad_arg<SMAG>(P, q, t, thv, Gm);
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
        int fnd;
        const int f = node_face_dir<Nq>(i, j, k, dir, fnd);
        if (f < 0) continue;
        const int64_t idx = fnd + (int64_t)Nfp * (f + 6 * e);
        const double *sgp = A.sg + 5 * idx;
        double nv[3] = {sgp[0], sgp[1], sgp[2]};
        const double sM = sgp[3], vMI = sgp[4];
        double Gs[3][5], nG[3][5], Gp[5];
        if (A.bnd[f + 6 * e] == 0) {
            const int32_t nb = A.nbr[idx];
            const int64_t ep = nb / Np;
            const int vp = nb - (int)ep * Np;
            double qp[5];
            const double *qpb = Q + (int64_t)Np * 5 * ep + vp;
#pragma unroll
            for (int s = 0; s < 5; ++s) qp[s] = qpb[Np * s];
            double Php = P.a_Phi >= 0 ? aux[(int64_t)Np * P.naux * ep + vp + (int64_t)Np * P.a_Phi] : -0.0;
            Thermo tp = thermo(P, qp, Php);
            double thp = SMAG ? theta_v(P, tp) : 0.0;
            grad_arg<SMAG>(P, qp, tp, thp, Gp);
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int s = 0; s < 5; ++s) Gs[d][s] = nv[d] * (Gp[s] + Gm[s]) * 0.5;
        } else {
            double qb[5];
            wall_state(q, nv, 1.0, qb);
            Thermo tb = thermo(P, qb, a.Phi);
            double thb = SMAG ? theta_v(P, tb) : 0.0;
            grad_arg<SMAG>(P, qb, tb, thb, Gp);
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int s = 0; s < 5; ++s) Gs[d][s] = nv[d] * Gp[s];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int s = 0; s < 5; ++s) nG[d][s] = nv[d] * Gm[s];
        double gfs[10], gfm[10];
        grad_flux<SMAG>(Gs, a, thv, gfs);
        grad_flux<SMAG>(nG, a, thv, gfm);
        const double w = vMI * sM;
#pragma unroll
        for (int s = 0; s < NGF; ++s) gf[s] += w * (gfs[s] - gfm[s]);
    }
    double *go = gfo + (int64_t)Np * NGF * e + n;
#pragma unroll
    for (int s = 0; s < NGF; ++s) go[Np * s] = gf[s];
}

// ------------------------------------------------------------------ tendency kernel
// MODE 0: tend = alpha*RHS + beta*tend.   MODE 1: LSRK stage, tend is dQ:
//   dQ <- RHS + dQ; Qout = Qin + bdt*dQ; dQ *= rka.  In MODE 1 beta = 0 says dQ holds
//   zeros (the previous stage scaled it by rka = 0) and is not read; dq_store = 0 leaves
//   the scaled dQ unstored (this stage's rka is 0 and the next stage gets beta = 0).
template <int Nq, bool DIFF, bool SMAG, int MODE>
__global__ void __launch_bounds__(threads_for<Nq>())
k_tendency(KParams P, KArrays A, const double *__restrict__ Q, const double *__restrict__ gfi,
           double *__restrict__ aux, double *__restrict__ tend, double alpha, double beta,
           double *__restrict__ Qout, double rka, double bdt, int write_aux, int dq_store) {
    constexpr int Np = Nq * Nq * Nq, Nfp = Nq * Nq;
    constexpr int NGF = SMAG ? 10 : 9;
    __shared__ double sD[Nq * Nq];
    __shared__ double sF[3][5][Np];
    const int64_t e = xcd_element(blockIdx.x, A.nelem);
    if (e >= A.nelem) return;
    const int n = threadIdx.x;
    if (n < Nq * Nq) sD[n] = A.D[n];
    const bool act = n < Np;
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);

    double q[5], gf[10], F1[3][5], F2[3][5], src[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    Aux a;
    Thermo t;
    double MI = 0.0;
    if (act) {
        const double *qb = Q + (int64_t)Np * 5 * e + n;
#pragma unroll
        for (int s = 0; s < 5; ++s) q[s] = qb[Np * s];
        a = load_aux(P, aux, Np, n, e);
        t = thermo(P, q, a.Phi);
        if (write_aux) {
            double *ab = aux + (int64_t)Np * P.naux * e + n;
            if (P.a_theta_v >= 0) ab[(int64_t)Np * P.a_theta_v] = theta_v(P, t);
            if (P.a_air_T >= 0) ab[(int64_t)Np * P.a_air_T] = t.temp;
        }
        flux1(P, q, t.p, a.ref_p, F1);
        if (DIFF) {
            const double *gb = gfi + (int64_t)Np * NGF * e + n;
#pragma unroll
            for (int s = 0; s < NGF; ++s) gf[s] = gb[Np * s];
            flux2<SMAG>(P, q, gf, a, F2);
        }
        const double *vg = A.vg + (int64_t)Np * NVG * e + n;
        const double M = vg[Np * 9];
        MI = vg[Np * 10];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const double x1 = vg[Np * ax], x2 = vg[Np * (ax + 3)], x3 = vg[Np * (ax + 6)];
#pragma unroll
            for (int s = 0; s < 5; ++s) {
                double Fx = F1[0][s], Fy = F1[1][s], Fz = F1[2][s];
                if (DIFF) { Fx += F2[0][s]; Fy += F2[1][s]; Fz += F2[2][s]; }
                sF[ax][s][n] = M * ((x1 * Fx + x2 * Fy) + x3 * Fz);
            }
        }
        if (P.gravity) {
            // Gravity() source (tendencies_momentum.jl:60-72)
            double drho = P.ref_src ? (q[0] - a.ref_rho) : q[0];
#pragma unroll
            for (int d = 0; d < 3; ++d) src[1 + d] = (-drho) * a.gPhi[d];
        }
    }
    __syncthreads();
    if (!act) return;

    // horizontal pass (written as alpha*T + beta*tendency), then vertical pass with
    // the source added at its own level (DGModel_kernels.jl:281-306, 453-545)
    double *tb = tend + (int64_t)Np * 5 * e + n;
    double T[5];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        double th = 0.0;
#pragma unroll
        for (int m = 0; m < Nq; ++m) {
            th += MI * sD[m + Nq * i] * sF[0][s][m + Nq * (j + Nq * k)];
            th += MI * sD[m + Nq * j] * sF[1][s][i + Nq * (m + Nq * k)];
        }
        double out;
        if (MODE == 0) out = (beta != 0.0) ? alpha * th + beta * tb[Np * s] : alpha * th;
        else out = (beta != 0.0) ? th + tb[Np * s] : th;
        double tv = 0.0;
#pragma unroll
        for (int m = 0; m < Nq; ++m) {
            tv += MI * sD[m + Nq * k] * sF[2][s][i + Nq * (j + Nq * m)];
            if (m == k) tv += src[s];
        }
        T[s] = (MODE == 0 ? alpha * tv : tv) + out;
    }

    // face fluxes of the faces this node lies on (NumericalFluxes.jl:219-340, 670-715)
    const bool on_face = i == 0 || i == Nq - 1 || j == 0 || j == Nq - 1 || k == 0 || k == Nq - 1;
    double csm = 0.0;
    if (on_face && P.rusanov) csm = sqrt(P.gammaR * t.temp);
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
        int fnd;
        const int f = node_face_dir<Nq>(i, j, k, dir, fnd);
        if (f < 0) continue;
        const int64_t idx = fnd + (int64_t)Nfp * (f + 6 * e);
        const double *sgp = A.sg + 5 * idx;
        const double nv[3] = {sgp[0], sgp[1], sgp[2]};
        const double w = (MODE == 0 ? alpha * sgp[4] : sgp[4]) * sgp[3];  // alpha * vMI * sM
        const bool wall = A.bnd[f + 6 * e] != 0;
        double qp[5], Fp[3][5], Phip, refpp;
        Aux ap;
        if (!wall) {
            const int32_t nb = A.nbr[idx];
            const int64_t ep = nb / Np;
            const int vp = nb - (int)ep * Np;
            const double *qpb = Q + (int64_t)Np * 5 * ep + vp;
#pragma unroll
            for (int s = 0; s < 5; ++s) qp[s] = qpb[Np * s];
            ap = load_aux(P, aux, Np, vp, ep);
            Phip = ap.Phi;
            refpp = ap.ref_p;
        } else {
            wall_state(q, nv, 2.0, qp);
            ap = a;
            Phip = a.Phi;
            refpp = a.ref_p;
        }
        Thermo tp = thermo(P, qp, Phip);
        flux1(P, qp, tp.p, refpp, Fp);
        const double h[3] = {nv[0] * 0.5, nv[1] * 0.5, nv[2] * 0.5};
        double flux[5];
#pragma unroll
        for (int s = 0; s < 5; ++s)
            flux[s] = -0.0 + (((F1[0][s] + Fp[0][s]) * h[0] + (F1[1][s] + Fp[1][s]) * h[1]) + (F1[2][s] + Fp[2][s]) * h[2]);
        if (P.rusanov) {
            const double rim = 1.0 / q[0], rip = 1.0 / qp[0];
            const double uNm = fabs((nv[0] * (rim * q[1]) + nv[1] * (rim * q[2])) + nv[2] * (rim * q[3]));
            const double uNp = fabs((nv[0] * (rip * qp[1]) + nv[1] * (rip * qp[2])) + nv[2] * (rip * qp[3]));
            const double mw = fmax(uNm + csm, uNp + sqrt(P.gammaR * tp.temp));
#pragma unroll
            for (int s = 0; s < 5; ++s) flux[s] += (mw * (q[s] - qp[s])) * 0.5;
        }
        if (DIFF && !wall) {
            const int32_t nb = A.nbr[idx];
            const int64_t ep = nb / Np;
            const int vp = nb - (int)ep * Np;
            double gfp[10];
            const double *gpb = gfi + (int64_t)Np * NGF * ep + vp;
#pragma unroll
            for (int s = 0; s < NGF; ++s) gfp[s] = gpb[Np * s];
            flux2<SMAG>(P, qp, gfp, ap, Fp);
#pragma unroll
            for (int s = 0; s < 5; ++s)
                flux[s] += ((F2[0][s] + Fp[0][s]) * h[0] + (F2[1][s] + Fp[1][s]) * h[1]) + (F2[2][s] + Fp[2][s]) * h[2];
        }
#pragma unroll
        for (int s = 0; s < 5; ++s) T[s] -= w * flux[s];
    }

    if (MODE == 0) {
#pragma unroll
        for (int s = 0; s < 5; ++s) tb[Np * s] = T[s];
    } else {
        double *qo = Qout + (int64_t)Np * 5 * e + n;
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            qo[Np * s] = q[s] + bdt * T[s];
            if (dq_store) tb[Np * s] = T[s] * rka;
        }
    }
}

}  // namespace cmdg
