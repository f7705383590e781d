// host_common.hpp -- host helpers shared by the single-block (api.hip) and the
// batch (batch.hip) entry points: descriptor checks, the geometry record the
// kernels read, and the translation of a control block into a mara_report.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/mara_gpu.h"
#include "kernels.hpp"
#include "fused.hpp"
#include "stage.hpp"

namespace mara {
namespace host {

inline bool hip_ok(hipError_t e)
{
    if (e != hipSuccess)
    {
        std::fprintf(stderr, "mara_gpu: HIP error %d (%s)\n", (int)e, hipGetErrorString(e));
        return false;
    }
    return true;
}

inline long long round_up(long long x, long long m) { return (x + m - 1) / m * m; }
// Law codes of the kernels: the three of the ABI, and 3 for NewtonianMHD
// with a positive cooling rate (nine fields: the auxiliary one last).
constexpr int LAW_MHD_COOLING = 3;
inline int law_nq(int law) { return law == MARA_LAW_HYDRO ? 5 : law == LAW_MHD_COOLING ? 9 : 8; }
inline int law_code(const mara_desc& d)
{
    return d.law == MARA_LAW_MHD && d.cooling_rate > 0.0 ? LAW_MHD_COOLING : d.law;
}
inline int recon_ng(int recon) { return recon == MARA_RECON_PCM ? 1 : recon == MARA_RECON_PLM ? 2 : 3; }

// UnitVector::xhat/yhat/zhat cartesian components (src/UnitVector.cpp:9-39):
// (sin(acos mu) cos phi, sin(acos mu) sin phi, mu) with (mu,phi) = (0,0),
// (0, pi/2), (1, 0), evaluated with the host libm like the reference.
inline void unit_vectors(double out[9])
{
    const double mu[3] = {0.0, 0.0, 1.0};
    const double phi[3] = {0.0, M_PI / 2, 0.0};
    for (int a = 0; a < 3; ++a)
    {
        const double st = std::sqrt(1 - mu[a] * mu[a]);
        out[3 * a + 0] = st * std::cos(phi[a]);
        out[3 * a + 1] = st * std::sin(phi[a]);
        out[3 * a + 2] = mu[a];
    }
}

inline int validate(const mara_desc* d)
{
    if (!d) return MARA_ERR_INVALID_ARGUMENT;
    if (d->law < 0 || d->law > 2) return MARA_ERR_INVALID_ARGUMENT;
    if (d->recon < 0 || d->recon > 2) return MARA_ERR_INVALID_ARGUMENT;
    if (d->riemann < 0 || d->riemann > 1) return MARA_ERR_INVALID_ARGUMENT;
    if (d->bc < 0 || d->bc > 2) return MARA_ERR_INVALID_ARGUMENT;
    if (d->rk_order < 1 || d->rk_order > 3) return MARA_ERR_INVALID_ARGUMENT;
    if (d->precision < 0 || d->precision > 2) return MARA_ERR_INVALID_ARGUMENT;
    if (d->riemann == MARA_RIEMANN_HLLC && d->law != MARA_LAW_HYDRO) return MARA_ERR_UNSUPPORTED;
    // the cooling source is a NewtonianMHD option
    if (d->cooling_rate > 0.0 && d->law != MARA_LAW_MHD) return MARA_ERR_UNSUPPORTED;
    const int ng = recon_ng(d->recon);
    for (int a = 0; a < 3; ++a)
    {
        if (d->cells[a] < 1 || d->guards[a] < 0) return MARA_ERR_INVALID_ARGUMENT;
        if (!(d->upper[a] > d->lower[a])) return MARA_ERR_INVALID_ARGUMENT;
        const int N = d->cells[a] + 2 * d->guards[a];
        // SchemeHelpers::makeFootprint (src/SolutionSchemes.cpp:45-48)
        if (N > 1 && d->guards[a] < ng) return MARA_ERR_GUARD_TOO_SMALL;
        // the ghost fill reads only interior zones
        if (d->guards[a] > 0 && d->cells[a] < d->guards[a]) return MARA_ERR_UNSUPPORTED;
    }
    if (d->cells[0] == 1 && (d->cells[1] > 1 || d->cells[2] > 1)) return MARA_ERR_UNSUPPORTED;
    if (d->geometry < MARA_GEOM_CARTESIAN || d->geometry > MARA_GEOM_SPHERICAL) return MARA_ERR_INVALID_ARGUMENT;
    if (d->geometry == MARA_GEOM_SPHERICAL)
    {
        // CellCenteredFieldCT is the Cartesian construction
        if (d->law != MARA_LAW_HYDRO && d->field_ct) return MARA_ERR_UNSUPPORTED;
        if (d->log_radial && !(d->lower[0] > 0.0)) return MARA_ERR_INVALID_ARGUMENT;
    }
    return MARA_OK;
}

// SphericalMeshGeometry::getEdge (src/SphericalMeshGeometry.cpp:190-205):
// edge n (interior-relative, fractional for centres) along an axis.
inline double sph_edge(const mara_desc& d, double n, int axis)
{
    if (axis == 0 && d.log_radial) return d.lower[0] * std::pow(d.upper[0] / d.lower[0], n / d.cells[0]);
    return d.lower[axis] + (d.upper[axis] - d.lower[axis]) * n / d.cells[axis];
}

// The tables of SphView for a spherical block (host libm), in its layout.
inline std::vector<double> spherical_tables(const mara_desc& d, const Geom& G)
{
    std::vector<double> t;
    t.reserve((size_t)SphView::size(G));
    const int* g = G.g;
    for (int n = 0; n <= G.N[0]; ++n) t.push_back(sph_edge(d, n - g[0], 0));
    for (int n = 0; n <= G.N[1]; ++n) t.push_back(sph_edge(d, n - g[1], 1));
    for (int n = 0; n <= G.N[1]; ++n) t.push_back(std::cos(sph_edge(d, n - g[1], 1)));
    for (int n = 0; n <= G.N[1]; ++n) t.push_back(std::sin(sph_edge(d, n - g[1], 1)));
    for (int n = 0; n <= G.N[2]; ++n) t.push_back(sph_edge(d, n - g[2], 2));
    // cellLength's radius (src/SphericalMeshGeometry.cpp:79-81); the linear
    // case takes its second edge from axis 1, as the reference does
    for (int i = 0; i < G.N[0]; ++i)
    {
        const int ic = i - g[0];
        t.push_back(d.log_radial ? std::sqrt(sph_edge(d, ic, 0) * sph_edge(d, ic + 1, 0))
                                 : 0.5 * (sph_edge(d, ic, 0) + sph_edge(d, ic + 1, 1)));
    }
    for (int j = 0; j < G.N[1]; ++j)
    {
        const int jc = j - g[1];
        t.push_back(std::sin(0.5 * (sph_edge(d, jc, 1) + sph_edge(d, jc + 1, 1))));
    }
    // coordinateAtIndex of the interior cells
    for (int a = 0; a < 3; ++a)
        for (int n = 0; n < G.n[a]; ++n) t.push_back(sph_edge(d, n + 0.5, a));
    return t;
}

// Extents, strides and the constants of CartesianMeshGeometry
// (src/CartesianMeshGeometry.cpp:145-157) and MeshOperator::linearCellDimension.
inline Geom make_geom(const mara_desc& d)
{
    Geom G;
    std::memset(&G, 0, sizeof G);
    for (int a = 0; a < 3; ++a)
    {
        G.N[a] = d.cells[a] + 2 * d.guards[a];
        G.g[a] = d.guards[a];
        G.n[a] = d.cells[a];
    }
    G.nq = law_nq(law_code(d));
    G.bc = d.bc;
    G.fs = round_up((long long)G.N[0] * G.N[1] * G.N[2], 64);
    G.ffs = round_up((long long)(G.N[0] + 1) * (G.N[1] + 1) * (G.N[2] + 1), 64);
    G.gamma = d.gamma;
    G.pfloor = d.pressure_floor;
    G.cool = law_code(d) == LAW_MHD_COOLING ? d.cooling_rate : 0.0;
    G.theta = d.plm_theta;
    double dx[3];
    for (int a = 0; a < 3; ++a) dx[a] = (d.upper[a] - d.lower[a]) / d.cells[a];
    G.dA[0] = dx[1] * dx[2];
    G.dA[1] = dx[2] * dx[0];
    G.dA[2] = dx[0] * dx[1];
    G.dV = dx[0] * dx[1] * dx[2];
    // std::min_element of the three lengths
    double L = dx[0];
    if (dx[1] < L) L = dx[1];
    if (dx[2] < L) L = dx[2];
    G.Lmin = L;
    unit_vectors(G.nhat);
    return G;
}

// status of a finished call, in the reference's order of precedence
// A launch shape whose values are all in the candidate lists.
inline bool valid_shape(const mara_shape& s)
{
    auto in = [](int v, const int* list, int n) {
        for (int i = 0; i < n; ++i)
            if (list[i] == v) return true;
        return false;
    };
    return in(s.pencil, PENCIL_SHAPES, 3) && s.zplanes >= 1 && s.zplanes <= KZ_MAX && in(s.rows, ROWS_SHAPES, 3);
}

inline int ctl_status(const Ctl& c)
{
    if (c.superluminal) return MARA_ERR_SUPERLUMINAL;
    if (c.eigen_fail) return MARA_ERR_BAD_EIGENVALUES;
    if (c.c2p_fail) return MARA_ERR_INVERSION_FAILURE;
    return MARA_OK;
}

inline int fill_report(const Ctl& c, const Geom& G, mara_report* r)
{
    const int status = ctl_status(c);
    if (r)
    {
        r->status = status;
        r->c2p_failures = (long long)c.c2p_fail;
        r->floor_applied = (long long)c.floor_applied;
        r->eigen_failures = (long long)c.eigen_fail;
        r->superluminal = (long long)c.superluminal;
        r->first_failure[0] = r->first_failure[1] = r->first_failure[2] = -1;
        if (c.c2p_fail)
        {
            const unsigned long long lin = c.first_fail;
            const unsigned long long n1 = G.n[1], n2 = G.n[2];
            r->first_failure[2] = (int)(lin % n2);
            r->first_failure[1] = (int)((lin / n2) % n1);
            r->first_failure[0] = (int)(lin / (n1 * n2));
        }
    }
    return status;
}

} // namespace host
} // namespace mara
