// newell.cpp -- host FP64 demag tensor (Newell-Williams-Dunlop, with the
// point-dipole far field and macro-PBC image sums).
//
// Restates src/micro/demag.jl:133-311 and src/atomistic/demag.jl:106-140 of
// MicroMagnetic.jl.  The operation order of every formula is the reference's,
// and this file is compiled with -ffp-contract=off, so the tensor rounds the
// way the reference's CPU path does: the 27-point Newell combination cancels
// catastrophically at large distances, so the order is part of the result.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace mmx {

static double newell_f(double x, double y, double z)
{
    double x2 = x * x, y2 = y * y, z2 = z * z;
    double R = std::sqrt(x2 + y2 + z2);
    if (R == 0.0) return 0.0;
    double f = 1.0 / 6 * (2 * x2 - y2 - z2) * R;
    if (x2 > 0) f -= x * y * z * std::atan(y * z / (x * R));
    if (x2 + z2 > 0) f += 0.5 * y * (z2 - x2) * std::asinh(y / (std::sqrt(x2 + z2)));
    if (x2 + y2 > 0) f += 0.5 * z * (y2 - x2) * std::asinh(z / (std::sqrt(x2 + y2)));
    return f;
}

static double newell_g(double x, double y, double z)
{
    double x2 = x * x, y2 = y * y, z2 = z * z;
    double R = std::sqrt(x2 + y2 + z2);
    if (R == 0.0) return 0.0;
    double g = -1.0 / 3 * x * y * R;
    if (z2 > 0) g -= 1.0 / 6 * z2 * z * std::atan(x * y / (z * R));
    if (y2 > 0) g -= 0.5 * y2 * z * std::atan(x * z / (y * R));
    if (x2 > 0) g -= 0.5 * x2 * z * std::atan(y * z / (x * R));
    if (x2 + y2 > 0) g += x * y * z * std::asinh(z / (std::sqrt(x2 + y2)));
    if (y2 + z2 > 0) g += 1.0 / 6 * y * (3 * z2 - y2) * std::asinh(x / (std::sqrt(y2 + z2)));
    if (x2 + z2 > 0) g += 1.0 / 6 * x * (3 * z2 - x2) * std::asinh(y / (std::sqrt(x2 + z2)));
    return g;
}

template <class F>
static double combo27(F fn, double x, double y, double z, double dx, double dy, double dz)
{
    double t = 8.0 * fn(x, y, z);
    t -= 4.0 * fn(x + dx, y, z);
    t -= 4.0 * fn(x - dx, y, z);
    t -= 4.0 * fn(x, y - dy, z);
    t -= 4.0 * fn(x, y + dy, z);
    t -= 4.0 * fn(x, y, z - dz);
    t -= 4.0 * fn(x, y, z + dz);
    t += 2.0 * fn(x + dx, y + dy, z);
    t += 2.0 * fn(x + dx, y - dy, z);
    t += 2.0 * fn(x - dx, y - dy, z);
    t += 2.0 * fn(x - dx, y + dy, z);
    t += 2.0 * fn(x + dx, y, z + dz);
    t += 2.0 * fn(x + dx, y, z - dz);
    t += 2.0 * fn(x - dx, y, z + dz);
    t += 2.0 * fn(x - dx, y, z - dz);
    t += 2.0 * fn(x, y - dy, z - dz);
    t += 2.0 * fn(x, y - dy, z + dz);
    t += 2.0 * fn(x, y + dy, z + dz);
    t += 2.0 * fn(x, y + dy, z - dz);
    t -= fn(x + dx, y + dy, z + dz);
    t -= fn(x + dx, y + dy, z - dz);
    t -= fn(x + dx, y - dy, z + dz);
    t -= fn(x + dx, y - dy, z - dz);
    t -= fn(x - dx, y + dy, z + dz);
    t -= fn(x - dx, y + dy, z - dz);
    t -= fn(x - dx, y - dy, z + dz);
    t -= fn(x - dx, y - dy, z - dz);
    return t / (4.0 * M_PI * dx * dy * dz);
}

static double dip_xx(double x, double y, double z)
{
    double R = x * x + y * y + z * z;
    if (R == 0) return 0.0;
    return -(2 * x * x - y * y - z * z) / (R * R * std::sqrt(R));
}

static double dip_xy(double x, double y, double z)
{
    double R = x * x + y * y + z * z;
    if (R == 0) return 0.0;
    return -3 * x * y / (R * R * std::sqrt(R));
}

static double max3(double a, double b, double c) { return std::max(std::max(a, b), c); }

static double txx(double x, double y, double z, double dx, double dy, double dz)
{
    double R = std::sqrt(x * x + y * y + z * z) / max3(dx, dy, dz);
    if (R > 60) return dip_xx(x, y, z) * dx * dy * dz / (4 * M_PI);
    return combo27(newell_f, x, y, z, dx, dy, dz);
}

static double txy(double x, double y, double z, double dx, double dy, double dz)
{
    double R = std::sqrt(x * x + y * y + z * z) / max3(dx, dy, dz);
    if (R > 60) return dip_xy(x, y, z) * dx * dy * dz / (4 * M_PI);
    return combo27(newell_g, x, y, z, dx, dy, dz);
}

static double tcomp(int c, double x, double y, double z, double dx, double dy, double dz)
{
    switch (c) {
    case 0: return txx(x, y, z, dx, dy, dz);
    case 1: return txx(y, x, z, dy, dx, dz);
    case 2: return txx(z, y, x, dz, dy, dx);
    case 3: return txy(x, y, z, dx, dy, dz);
    case 4: return txy(x, z, y, dx, dz, dy);
    default: return txy(y, z, x, dy, dz, dx);
    }
}

// Compact tensor out[c*N + (k*ny + j)*nx + i], c = xx,yy,zz,xy,xz,yz, for
// displacements (i,j,k) >= 0 in units of the largest cell edge, summed over
// the (2Nx+1)(2Ny+1)(2Nz+1) images (tensors_kernel_*!, demag.jl:298-386).
void newell_tensor(int nx, int ny, int nz, double Dx, double Dy, double Dz, int Nx, int Ny, int Nz,
                   double *out, int nthreads)
{
    const double ms = max3(Dx, Dy, Dz);
    const double dx = Dx / ms, dy = Dy / ms, dz = Dz / ms;
    const int64_t N = (int64_t)nx * ny * nz;
    const int64_t rows = (int64_t)ny * nz * 6;
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    nthreads = (int)std::min<int64_t>(nthreads, std::max<int64_t>(1, rows));
    auto work = [&](int tid) {
        for (int64_t r = tid; r < rows; r += nthreads) {
            int c = (int)(r / ((int64_t)ny * nz));
            int64_t q = r % ((int64_t)ny * nz);
            int j = (int)(q % ny), k = (int)(q / ny);
            double *o = out + c * N + ((int64_t)k * ny + j) * nx;
            for (int i = 0; i < nx; i++) {
                double sum = 0.0;
                for (int p = -Nx; p <= Nx; p++)
                    for (int qq = -Ny; qq <= Ny; qq++)
                        for (int s = -Nz; s <= Nz; s++) {
                            double x = (i + (double)p * nx) * dx;
                            double y = (j + (double)qq * ny) * dy;
                            double z = (k + (double)s * nz) * dz;
                            sum += tcomp(c, x, y, z, dx, dy, dz);
                        }
                o[i] = sum;
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nthreads; t++) th.emplace_back(work, t);
    work(0);
    for (auto &t : th) t.join();
}

}  // namespace mmx
