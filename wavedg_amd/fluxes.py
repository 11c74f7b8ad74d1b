"""Flux functors as HIP source for capi.Flux (DivF2D<true> /
EdgeFluxF2D<true> with functors compiled at run time).  Each string defines
the functors in the reference's signatures; the library appends its kernels.
"""

# examples/cl2d.cpp: the compressible Euler flux F (:57-81) and the local
# Lax-Friedrichs numerical flux (speed :84-96, numerical_flux :98-111), for
# q = (rho, rho u, rho v, E), n_var = 4.
EULER_LLF = r"""
#define GAMMA 1.4
__device__ inline double euler_speed(const double* q)
{
    const double rho = q[0], u = q[1] / rho, v = q[2] / rho;
    const double p = (GAMMA - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v));
    return hypot(u, v) + sqrt(GAMMA * p / rho);
}
__device__ inline void euler_fg(const double* q, double* f, double* g)
{
    const double rho = q[0], u = q[1] / rho, v = q[2] / rho, E = q[3];
    const double p = (GAMMA - 1.0) * (E - 0.5 * rho * (u * u + v * v));
    f[0] = rho * u; f[1] = rho * u * u + p; f[2] = rho * u * v; f[3] = u * (E + p);
    g[0] = rho * v; g[1] = rho * u * v; g[2] = rho * v * v + p; g[3] = v * (E + p);
}
__device__ void flux(const double* x, const double* q, double* F)
{
    euler_fg(q, F, F + 4);
}
__device__ void numerical_flux(const double* x, const double* n, const double* qL, const double* qR, double* fh)
{
    double fL[4], gL[4], fR[4], gR[4];
    euler_fg(qL, fL, gL);
    euler_fg(qR, fR, gR);
    const double cL = euler_speed(qL), cR = euler_speed(qR), c = cL > cR ? cL : cR;
    for (int d = 0; d < 4; ++d)
        fh[d] = 0.5 * n[0] * (fL[d] + fR[d]) + 0.5 * n[1] * (gL[d] + gR[d]) + 0.5 * c * (qL[d] - qR[d]);
}
"""

# The wave equation's upwind flux a/2 C (uL + uR) + b C^2 (uL - uR), C = n.A,
# a = -1, b = -1/2 (EdgeFlux<true> of WaveEquation, EdgeFlux.cpp:31-139 with
# |C| = C^2 for unit n), the reflecting exterior state (WaveEquation.hpp:64-69)
# and the wave fluxes f = (u, p, 0), g = (v, 0, p); n_var = 3.
WAVE_UPWIND = r"""
#define WDG_EXTERIOR_STATE
__device__ void flux(const double* x, const double* w, double* F)
{
    F[0] = w[1]; F[1] = w[0]; F[2] = 0.0;
    F[3] = w[2]; F[4] = 0.0;  F[5] = w[0];
}
__device__ void numerical_flux(const double* x, const double* n, const double* uL, const double* uR, double* fh)
{
    const double sp = uL[0] + uR[0], su = uL[1] + uR[1], sv = uL[2] + uR[2];
    const double dp = uL[0] - uR[0], du = uL[1] - uR[1], dv = uL[2] - uR[2];
    const double ndu = n[0] * du + n[1] * dv;
    fh[0] = -0.5 * (n[0] * su + n[1] * sv) - 0.5 * dp;
    fh[1] = -0.5 * n[0] * sp - 0.5 * n[0] * ndu;
    fh[2] = -0.5 * n[1] * sp - 0.5 * n[1] * ndu;
}
__device__ void exterior_state(const double* x, const double* n, const double* u, double* e)
{
    e[0] = u[0];
    e[1] = (n[1] * n[1] - n[0] * n[0]) * u[1] - 2.0 * n[0] * n[1] * u[2];
    e[2] = -2.0 * n[0] * n[1] * u[1] + (n[0] * n[0] - n[1] * n[1]) * u[2];
}
"""
