// generators.cpp -- deterministic SPD workloads (host C++).
#include "generators.hpp"

#include <cmath>
#include <stdexcept>

namespace xp {

double Rng::normal() {
    double u1 = uniform();
    const double u2 = uniform();
    if (u1 < 0x1.0p-60) u1 = 0x1.0p-60;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925 * u2);
}

uint64_t stream_seed(uint64_t seed, uint64_t k) {
    Rng r(seed ^ (0xD1B54A32D192ED03ull * (k + 1)));
    return r.next();
}

namespace {

void fill_rhs(HostCsr& a, uint64_t seed) {
    Rng rb(stream_seed(seed, 99));
    a.b.resize(static_cast<size_t>(a.n));
    for (int i = 0; i < a.n; ++i) a.b[static_cast<size_t>(i)] = rb.uniform(-1.0, 1.0);
    a.x0.assign(static_cast<size_t>(a.n), 0.0);
}

}  // namespace

HostCsr gen_diffusion2d(int nx, int ny, double sigma, uint64_t seed) {
    if (nx <= 0 || ny <= 0) throw std::runtime_error("gen_diffusion2d: bad grid");
    Rng rk(stream_seed(seed, 1));
    // Edge conductivities: x-edges (nx+1 per grid row, the two ends are
    // Dirichlet ghost edges), then y-edges (ny+1 per grid column).
    std::vector<double> kx(static_cast<size_t>(nx + 1) * ny), ky(static_cast<size_t>(ny + 1) * nx);
    for (auto& v : kx) v = std::exp(sigma * rk.normal());
    for (auto& v : ky) v = std::exp(sigma * rk.normal());
    auto KX = [&](int i, int j) { return kx[static_cast<size_t>(j) * (nx + 1) + i]; };  // edge left of (i,j)
    auto KY = [&](int i, int j) { return ky[static_cast<size_t>(i) * (ny + 1) + j]; };  // edge below (i,j)
    HostCsr a;
    a.n = nx * ny;
    a.row_ptr.reserve(static_cast<size_t>(a.n) + 1);
    a.row_ptr.push_back(0);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const double w = KX(i, j), e = KX(i + 1, j), s = KY(i, j), nn = KY(i, j + 1);
            const int id = j * nx + i;
            if (j > 0) { a.col_idx.push_back(id - nx); a.values.push_back(-s); }
            if (i > 0) { a.col_idx.push_back(id - 1); a.values.push_back(-w); }
            a.col_idx.push_back(id);
            a.values.push_back(w + e + s + nn);
            if (i + 1 < nx) { a.col_idx.push_back(id + 1); a.values.push_back(-e); }
            if (j + 1 < ny) { a.col_idx.push_back(id + nx); a.values.push_back(-nn); }
            a.row_ptr.push_back(static_cast<int>(a.col_idx.size()));
        }
    fill_rhs(a, seed);
    return a;
}

HostCsr gen_diffusion3d(int k, double sigma, uint64_t seed) {
    if (k <= 0 || static_cast<long long>(k) * k * k > 200000000LL) throw std::runtime_error("gen_diffusion3d: bad grid");
    Rng rk(stream_seed(seed, 1));
    const size_t line = static_cast<size_t>(k) + 1, plane = static_cast<size_t>(k) * k;
    std::vector<double> ex(plane * line), ey(plane * line), ez(plane * line);
    for (auto& v : ex) v = std::exp(sigma * rk.normal());
    for (auto& v : ey) v = std::exp(sigma * rk.normal());
    for (auto& v : ez) v = std::exp(sigma * rk.normal());
    // ex[(z*k + y)*(k+1) + x] is the x-edge on the low-x side of node (x,y,z).
    auto EX = [&](int x, int y, int z) { return ex[(static_cast<size_t>(z) * k + y) * line + x]; };
    auto EY = [&](int x, int y, int z) { return ey[(static_cast<size_t>(z) * k + x) * line + y]; };
    auto EZ = [&](int x, int y, int z) { return ez[(static_cast<size_t>(y) * k + x) * line + z]; };
    HostCsr a;
    a.n = k * k * k;
    a.row_ptr.reserve(static_cast<size_t>(a.n) + 1);
    a.col_idx.reserve(static_cast<size_t>(a.n) * 7);
    a.values.reserve(static_cast<size_t>(a.n) * 7);
    a.row_ptr.push_back(0);
    const int kk = k * k;
    for (int z = 0; z < k; ++z)
        for (int y = 0; y < k; ++y)
            for (int x = 0; x < k; ++x) {
                const int id = (z * k + y) * k + x;
                const double xl = EX(x, y, z), xh = EX(x + 1, y, z);
                const double yl = EY(x, y, z), yh = EY(x, y + 1, z);
                const double zl = EZ(x, y, z), zh = EZ(x, y, z + 1);
                if (z > 0) { a.col_idx.push_back(id - kk); a.values.push_back(-zl); }
                if (y > 0) { a.col_idx.push_back(id - k); a.values.push_back(-yl); }
                if (x > 0) { a.col_idx.push_back(id - 1); a.values.push_back(-xl); }
                a.col_idx.push_back(id);
                a.values.push_back(xl + xh + yl + yh + zl + zh);
                if (x + 1 < k) { a.col_idx.push_back(id + 1); a.values.push_back(-xh); }
                if (y + 1 < k) { a.col_idx.push_back(id + k); a.values.push_back(-yh); }
                if (z + 1 < k) { a.col_idx.push_back(id + kk); a.values.push_back(-zh); }
                a.row_ptr.push_back(static_cast<int>(a.col_idx.size()));
            }
    fill_rhs(a, seed);
    return a;
}

HostCsr gen_fem2d(int m, double sigma, int n_incl, double contrast, double jitter, uint64_t seed) {
    if (m < 3 || m > 20000) throw std::runtime_error("gen_fem2d: bad grid");
    const double h = 1.0 / (m - 1);
    // Inclusions first, so their positions do not depend on the grid size.
    Rng ri(stream_seed(seed, 2));
    std::vector<double> cx(static_cast<size_t>(n_incl)), cy(cx.size()), cr(cx.size());
    for (int q = 0; q < n_incl; ++q) {
        cx[static_cast<size_t>(q)] = ri.uniform();
        cy[static_cast<size_t>(q)] = ri.uniform();
        cr[static_cast<size_t>(q)] = ri.uniform(0.02, 0.06);
    }
    // Node positions, row-major over (iy, ix); boundary nodes stay on the square.
    Rng rj(stream_seed(seed, 3));
    std::vector<double> px(static_cast<size_t>(m) * m), py(px.size());
    for (int iy = 0; iy < m; ++iy)
        for (int ix = 0; ix < m; ++ix) {
            const size_t id = static_cast<size_t>(iy) * m + ix;
            px[id] = ix * h;
            py[id] = iy * h;
            if (ix > 0 && iy > 0 && ix < m - 1 && iy < m - 1) {
                px[id] += rj.uniform(-jitter * h, jitter * h);
                py[id] += rj.uniform(-jitter * h, jitter * h);
            }
        }
    Rng rd(stream_seed(seed, 4));
    Rng rk(stream_seed(seed, 5));
    const int mi = m - 2;  // interior nodes per direction
    // 3x3 stencil accumulator per interior node: slot (dy+1)*3 + (dx+1).
    std::vector<double> st(static_cast<size_t>(mi) * mi * 9, 0.0);
    std::vector<unsigned char> present(st.size(), 0);
    auto interior = [&](int ix, int iy) { return ix > 0 && iy > 0 && ix < m - 1 && iy < m - 1; };
    auto add_tri = [&](const int (&vx)[3], const int (&vy)[3], double kc) {
        double X[3], Y[3];
        for (int a = 0; a < 3; ++a) {
            const size_t id = static_cast<size_t>(vy[a]) * m + vx[a];
            X[a] = px[id];
            Y[a] = py[id];
        }
        const double det = (X[1] - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[1] - Y[0]);
        const double area = 0.5 * std::fabs(det);
        const double bb[3] = {Y[1] - Y[2], Y[2] - Y[0], Y[0] - Y[1]};
        const double cc[3] = {X[2] - X[1], X[0] - X[2], X[1] - X[0]};
        for (int a = 0; a < 3; ++a) {
            if (!interior(vx[a], vy[a])) continue;
            const size_t row = static_cast<size_t>(vy[a] - 1) * mi + (vx[a] - 1);
            for (int c = 0; c < 3; ++c) {
                if (!interior(vx[c], vy[c])) continue;
                const int slot = (vy[c] - vy[a] + 1) * 3 + (vx[c] - vx[a] + 1);
                st[row * 9 + static_cast<size_t>(slot)] += kc * (bb[a] * bb[c] + cc[a] * cc[c]) / (4.0 * area);
                present[row * 9 + static_cast<size_t>(slot)] = 1;
            }
        }
    };
    auto kappa = [&](const int (&vx)[3], const int (&vy)[3]) {
        double k = std::exp(sigma * rk.normal());
        double gx = 0.0, gy = 0.0;
        for (int a = 0; a < 3; ++a) {
            const size_t id = static_cast<size_t>(vy[a]) * m + vx[a];
            gx += px[id] / 3.0;
            gy += py[id] / 3.0;
        }
        for (int q = 0; q < n_incl; ++q) {
            const double dx = gx - cx[static_cast<size_t>(q)], dy = gy - cy[static_cast<size_t>(q)];
            if (dx * dx + dy * dy < cr[static_cast<size_t>(q)] * cr[static_cast<size_t>(q)]) {
                k *= contrast;
                break;
            }
        }
        return k;
    };
    for (int iy = 0; iy < m - 1; ++iy)
        for (int ix = 0; ix < m - 1; ++ix) {
            const bool main_diag = rd.uniform() < 0.5;  // split along (ix,iy)-(ix+1,iy+1)
            if (main_diag) {
                const int ax[3] = {ix, ix + 1, ix + 1}, ay[3] = {iy, iy, iy + 1};
                const int bx[3] = {ix, ix + 1, ix}, by[3] = {iy, iy + 1, iy + 1};
                const double ka = kappa(ax, ay), kb = kappa(bx, by);
                add_tri(ax, ay, ka);
                add_tri(bx, by, kb);
            } else {
                const int ax[3] = {ix, ix + 1, ix}, ay[3] = {iy, iy, iy + 1};
                const int bx[3] = {ix + 1, ix + 1, ix}, by[3] = {iy, iy + 1, iy + 1};
                const double ka = kappa(ax, ay), kb = kappa(bx, by);
                add_tri(ax, ay, ka);
                add_tri(bx, by, kb);
            }
        }
    HostCsr a;
    a.n = mi * mi;
    a.row_ptr.reserve(static_cast<size_t>(a.n) + 1);
    a.col_idx.reserve(static_cast<size_t>(a.n) * 7);
    a.values.reserve(static_cast<size_t>(a.n) * 7);
    a.row_ptr.push_back(0);
    for (int iy = 0; iy < mi; ++iy)
        for (int ix = 0; ix < mi; ++ix) {
            const size_t row = static_cast<size_t>(iy) * mi + ix;
            for (int slot = 0; slot < 9; ++slot) {
                if (!present[row * 9 + static_cast<size_t>(slot)]) continue;
                const int dy = slot / 3 - 1, dx = slot % 3 - 1;
                a.col_idx.push_back((iy + dy) * mi + (ix + dx));
                a.values.push_back(st[row * 9 + static_cast<size_t>(slot)]);
            }
            a.row_ptr.push_back(static_cast<int>(a.col_idx.size()));
        }
    fill_rhs(a, seed);
    return a;
}

}  // namespace xp
