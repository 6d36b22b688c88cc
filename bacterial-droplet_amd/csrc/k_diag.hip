// k_diag.hip -- on-device diagnostics of a droplet: the phase-field centroid
// and the centroid z-slice with its 4th-order vorticities.
//
// They replace the post-processing reads of full-field text snapshots:
//   readDraw.py:30-52 / characteristicVelocity.py:16-39 calculate_centroid:
//     sum of phi*i, phi*j, phi*k over the array axes (i, j, k) = (y, x, z) of
//     the snapshot's reshape (Ly, Lx, Lz), divided by sum phi, truncated to int;
//   readDraw.py:220-237: the z = int(centroid_z) slice of phi, v, P and the
//     vorticities d1xO4(vy) - d1yO4(vx), d1xO4(py) - d1yO4(px), with op.py:3-13
//     d1xO4 u = (u[x-2] - 8 u[x-1] + 8 u[x+1] - u[x+2]) / (12 h), periodic.
// The sums are FP64, accumulated per work-group in a fixed order.
#include "ttsh_dev.h"
#include "ttsh_kernels.h"

namespace ttsh {

// Per work-group FP64 partial sums of phi, phi*y, phi*x, phi*z.
__global__ __launch_bounds__(256) void k_centroid_part(const float* __restrict__ phi, int N, double* __restrict__ part,
                                                       int nblk)
{
    __shared__ double red[4 * 4];
    const size_t n = (size_t)N * N * N;
    double s0 = 0.0, sy = 0.0, sx = 0.0, sz = 0.0;
    const gptr<const float> f = glb(phi);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)nblk * blockDim.x) {
        const double p = f[i];
        const size_t x = i % N, y = (i / N) % N, z = i / ((size_t)N * N);
        s0 += p;
        sy += p * (double)y;
        sx += p * (double)x;
        sz += p * (double)z;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_down(s0, off, 64);
        sy += __shfl_down(sy, off, 64);
        sx += __shfl_down(sx, off, 64);
        sz += __shfl_down(sz, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red[4 * wid] = s0;
        red[4 * wid + 1] = sy;
        red[4 * wid + 2] = sx;
        red[4 * wid + 3] = sz;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double a = 0.0;
        for (int w = 0; w < 4; ++w) a += red[4 * w + threadIdx.x];
        glb(part)[4 * (size_t)blockIdx.x + threadIdx.x] = a;
    }
}

__global__ __launch_bounds__(64) void k_centroid_final(const double* __restrict__ part, int nblk, double* __restrict__ out)
{
    // one thread per sum, blocks in index order
    if (threadIdx.x < 4) {
        double a = 0.0;
        for (int b = 0; b < nblk; ++b) a += glb(part)[4 * (size_t)b + threadIdx.x];
        glb(out)[threadIdx.x] = a;
    }
}

// Plane z of phi, v, P and the two vorticities, [field][y][x] in out.  With
// zfix < 0 the plane is the centroid plane int(sum phi*z / sum phi) from the
// device sums (clamped to the box), and it is stored at out[9*N*N].
__global__ __launch_bounds__(256) void k_slice(const float* __restrict__ phi, const float* __restrict__ vx,
                                               const float* __restrict__ vy, const float* __restrict__ vz,
                                               const float* __restrict__ px, const float* __restrict__ py,
                                               const float* __restrict__ pz, int N, int zfix,
                                               const double* __restrict__ sums, float inv12h, float* __restrict__ out)
{
    const int NN = N * N;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int z = zfix;
    if (z < 0) {
        const double s0 = glb(sums)[0], sz = glb(sums)[3];
        z = (int)(sz / s0);  // truncation toward zero, as Python int()
        z = z < 0 ? 0 : (z > N - 1 ? N - 1 : z);
        if (i == 0) glb(out)[9 * NN] = (float)z;
    }
    if (i >= NN) return;
    const int x = i % N, y = i / N;
    const size_t pl = (size_t)z * NN;
    auto at = [&](const float* f, int xx, int yy) {
        xx = (xx + N) % N;
        yy = (yy + N) % N;
        return glb(f)[pl + (size_t)yy * N + xx];
    };
    // d/dx along the row, d/dy along the column: 4th-order central, periodic
    auto dx4 = [&](const float* f) {
        return inv12h * ((at(f, x - 2, y) - 8.f * at(f, x - 1, y)) + (8.f * at(f, x + 1, y) - at(f, x + 2, y)));
    };
    auto dy4 = [&](const float* f) {
        return inv12h * ((at(f, x, y - 2) - 8.f * at(f, x, y - 1)) + (8.f * at(f, x, y + 1) - at(f, x, y + 2)));
    };
    const gptr<float> o = glb(out);
    o[0 * NN + i] = at(phi, x, y);
    o[1 * NN + i] = at(vx, x, y);
    o[2 * NN + i] = at(vy, x, y);
    o[3 * NN + i] = at(vz, x, y);
    o[4 * NN + i] = at(px, x, y);
    o[5 * NN + i] = at(py, x, y);
    o[6 * NN + i] = at(pz, x, y);
    o[7 * NN + i] = dx4(vy) - dy4(vx);
    o[8 * NN + i] = dx4(py) - dy4(px);
}

hipError_t launch_centroid(const float* phi, int N, double* part, int nblk, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_centroid_part, dim3(nblk), dim3(256), 0, st, phi, N, part, nblk);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_centroid_final, dim3(1), dim3(64), 0, st, part, nblk, out);
    return hipGetLastError();
}

hipError_t launch_slice(const float* phi, const float* const v[3], const float* const P[3], int N, int z,
                        const double* sums, float h, float* out, hipStream_t st)
{
    const int NN = N * N;
    hipLaunchKernelGGL(k_slice, dim3((NN + 255) / 256), dim3(256), 0, st, phi, v[0], v[1], v[2], P[0], P[1], P[2], N, z,
                       sums, 1.f / (12.f * h), out);
    return hipGetLastError();
}

}  // namespace ttsh
