// The receivers' first-order reading on the device: W, X, Y, Z of each listener from the
// pressure at its 13 points, read straight from the stepper's probe output in device memory
// (the reference's _first_order, src/acoustic_spaces/engine/wave.py:757-802).
//
// Per listener: W is the pressure at the listener; X, Y, Z are the Courant number at the
// listener times the running trapezoidal integral of the pressure's change over a cell along
// each axis: the fourth-order central difference over two cells either side, the
// second-order one where a point two cells away is missing, a one-sided one where a point a
// cell away is, nothing where both are. The trapezoid's gain correction, a spectral step,
// is left to the caller (an FFT).
//
// One thread per listener and channel walks its samples in order: each channel's integral
// is a sequential sum, summed in the same order and with the same operations as the host's
// NumPy reading (halves h_k = d_k / 2, I_k = I_(k-1) + (h_k + h_(k-1))), so the two agree
// bit for bit (the library is built without contraction). The work is a few thousand
// threads over the run's samples, far below the step's.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "asfdtd.h"

namespace {

constexpr uint32_t kPoints = 13;  // points a listener is read at: itself, +-1 and +-2 cells along each axis

enum Difference : int { kNone = 0, kFourth = 1, kCentral = 2, kUp = 3, kDown = 4 };

__global__ void first_order_kernel(const double *__restrict__ probes, uint64_t steps, uint32_t listeners,
                                   const uint8_t *__restrict__ found, const double *__restrict__ courant,
                                   double *__restrict__ out) {
    const uint32_t task = blockIdx.x * blockDim.x + threadIdx.x;
    if (task >= listeners * 4u) return;
    const uint32_t l = task / 4u, channel = task % 4u;
    const double *p = probes + static_cast<uint64_t>(kPoints) * l * steps;
    double *dst = out + (static_cast<uint64_t>(l) * 4u + channel) * steps;
    if (channel == 0) {
        for (uint64_t k = 0; k < steps; ++k) dst[k] = p[k];
        return;
    }
    const uint8_t *f = found + kPoints * l;
    const uint32_t up = 1 + 2 * (channel - 1), down = up + 1;
    int mode = kNone;
    if (f[up] && f[down]) {
        mode = (f[up + 6] && f[down + 6]) ? kFourth : kCentral;
    } else if (f[up]) {
        mode = kUp;
    } else if (f[down]) {
        mode = kDown;
    }
    const double c = courant[l];
    const double *pu = p + up * steps, *pd = p + down * steps, *pl = p;
    const double *puu = p + (up + 6) * steps, *pdd = p + (down + 6) * steps;
    double integral = 0.0, previous = 0.0;
    for (uint64_t k = 0; k < steps; ++k) {
        double d = 0.0;
        if (mode == kFourth) {
            const double near = 0.5 * (pu[k] - pd[k]);
            const double far = puu[k] - pdd[k];
            d = near * (4.0 / 3.0) - far / 12.0;
        } else if (mode == kCentral) {
            d = 0.5 * (pu[k] - pd[k]);
        } else if (mode == kUp) {
            d = pu[k] - pl[k];
        } else if (mode == kDown) {
            d = pl[k] - pd[k];
        }
        const double half = 0.5 * d;
        integral = integral + (half + previous);
        previous = half;
        dst[k] = mode == kNone ? 0.0 : c * integral;
    }
}

}  // namespace

extern "C" int as_first_order_device(const double *d_probes, size_t steps, size_t listeners, const uint8_t *d_found,
                                     const double *d_courant, double *d_out, void *stream) {
    if (listeners == 0 || steps == 0) return AS_OK;
    if (d_probes == nullptr || d_found == nullptr || d_courant == nullptr || d_out == nullptr) return AS_EINVAL;
    if (listeners > (1u << 28)) return AS_ERANGE;
    const uint32_t tasks = static_cast<uint32_t>(listeners) * 4u;
    const uint32_t block = 64;
    hipLaunchKernelGGL(first_order_kernel, dim3((tasks + block - 1) / block), dim3(block), 0,
                       static_cast<hipStream_t>(stream), d_probes, static_cast<uint64_t>(steps),
                       static_cast<uint32_t>(listeners), d_found, d_courant, d_out);
    return hipGetLastError() == hipSuccess ? AS_OK : AS_EDEVICE;
}
