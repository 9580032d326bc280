// Compiled functions (heyoka's cfunc, heyoka/cfunc.cpp:90-121) on the GPU: one
// lane per evaluation point, compiled with hipRTC after the generated prelude,
// which defines HYX_N (inputs), HYX_NPAR (>= 1), HYX_NOUT and
//   hyx_cfunc_eval(x, pv, t0, out, os): the outputs at one point.
// Layout: inputs [HYX_N][ld_in], pars [npar][ld_par], outputs [HYX_NOUT][ld_out],
// point index fastest (a batch's device-resident state is an input as it is).

extern "C" __global__ void __launch_bounds__(64) hyx_cfunc_k(const double *in, hyx_u64 ld_in, const double *pars,
                                                            hyx_u64 ld_par, unsigned npar, const double *time,
                                                            double *out, hyx_u64 ld_out, hyx_u64 n_points)
{
    const hyx_u64 i = (hyx_u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_points) return;
    double x[HYX_N];
    double pv[HYX_NPAR];
#pragma unroll
    for (int q = 0; q < HYX_N; ++q) x[q] = in[(hyx_u64)q * ld_in + i];
#pragma unroll
    for (int q = 0; q < HYX_NPAR; ++q) pv[q] = ((unsigned)q < npar && pars) ? pars[(hyx_u64)q * ld_par + i] : 0.0;
    const double t0 = time ? time[i] : 0.0;
    hyx_cfunc_eval(x, pv, t0, out + i, ld_out);
}
