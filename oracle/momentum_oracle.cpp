// Momentum side of the SIMPLE corrector: an FP64 CPU restatement of the
// reference algorithm, used ONLY as test infrastructure. Only tests/ and the
// smoke check load the library built from it; libcfdxhip.so never does.
//
// Restated (paths relative to the stevendaix/cfdx_dev sources):
//   compute_scalar_gradient_with_boundary    physics/finite_volume_transport.h:505-586
//   weighted_least_squares_gradient_robust   core/fvm/least_squares_gradient.h:77-205
//   compute_boundary_aware_wls_gradient      physics/finite_volume_transport.h:606-715
//   compute_nonorthogonal_diffusion_face_fluxes                            :717-874
//   assemble_scalar_equation (UPWIND, cell viscosity, no face conditions)  :882-1310
//   assemble_momentum_component (cell-viscosity overload)
//                                            physics/steady_incompressible_solver.h:3262-3311
//   relax_momentum_equation                  :3728-3770
//   rAU = 1 / diagonal (SIMPLE)              :7816-7836
//   build_hbya                               :6430-6465
//   velocity correction                      :9545-9557
//   make_rhie_chow_mass_flux (HBYA_PREDICTOR, w = 1/2, no skew correction)
//                                            :2063-2217
//
// The loops keep the reference's structure and floating-point operation order
// (global face order for every per-cell accumulation), so that, compiled
// without FMA contraction, results are bit-identical to oracle/_ref; the
// tests check exactly that.
//
// Conventions: geometry arrays are AoS xyz triplets (oracle_geometry);
// vector cell fields are SoA, 3 x n_cells; patch_kind 1 = EMPTY, 2 = SYMMETRY;
// scalar bc_type per patch 0 zero-gradient, 1 fixed value, 2 fixed gradient
// (bc_value holds the value or the gradient); velocity vbc_type per patch
// 0 zero-gradient, 1 fixed value (3 values per patch).

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <map>
#include <vector>

namespace {

using std::size_t;

struct P3 {
    double x = 0.0, y = 0.0, z = 0.0;
};
inline P3 padd(const P3& a, const P3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 psub(const P3& a, const P3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 pmul(const P3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double pdot(const P3& a, const P3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double pmag2(const P3& a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double pmag(const P3& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline P3 at3(const double* a, size_t i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }

struct MeshView {
    size_t nf, nc, n_patches;
    const uint32_t* owner;
    const int32_t* neigh;
    const uint32_t* coff;
    const uint32_t* cfaces;
    const double *fc, *fs, *cc, *vol;
    const uint32_t* patch;
    const uint8_t* kind;
    bool empty_face(size_t f) const { return patch[f] < n_patches && kind[patch[f]] == 1; }
    // resolve_scalar_boundary_condition (finite_volume_transport.h:271-318):
    // the patch's condition, zero-gradient when the face has no patch.
    uint8_t bc_of(size_t f, const uint8_t* bct) const { return patch[f] < n_patches ? bct[patch[f]] : 0; }
};

// boundary_normal_distance: finite_volume_transport.h:247-269.
bool boundary_normal_distance(const MeshView& m, size_t f, size_t c, double& out) {
    const P3 Sf = at3(m.fs, f);
    const double area = pmag(Sf);
    if (!(area > 0.0) || !std::isfinite(area)) return false;
    const P3 d = psub(at3(m.fc, f), at3(m.cc, c));
    const double dn = std::abs(pdot(d, Sf)) / area;
    if (!(dn > 0.0) || !std::isfinite(dn)) {
        const double v = m.vol[c];
        if (v > 0.0 && std::isfinite(v)) {
            const double dn2 = std::sqrt(v);
            if (dn2 > 0.0 && std::isfinite(dn2)) {
                out = dn2;
                return true;
            }
        }
        return false;
    }
    out = dn;
    return true;
}

// compute_scalar_gradient_with_boundary: finite_volume_transport.h:505-586.
bool gg_gradient(const MeshView& m, const uint8_t* bct, const double* bcv, const double* field, double* grad) {
    const size_t nc = m.nc;
    for (size_t c = 0; c < nc; ++c) {
        P3 sum{};
        for (uint32_t k = m.coff[c]; k < m.coff[c + 1]; ++k) {
            const size_t f = m.cfaces[k];
            const bool own = m.owner[f] == c;
            const P3 sf = own ? at3(m.fs, f) : pmul(at3(m.fs, f), -1.0);
            double fv = field[c];
            if (m.neigh[f] >= 0) {
                const size_t other = own ? size_t(m.neigh[f]) : size_t(m.owner[f]);
                const double dc = pmag(psub(at3(m.fc, f), at3(m.cc, c)));
                const double dn = pmag(psub(at3(m.cc, other), at3(m.fc, f)));
                if (!(dc + dn > 0.0) || !std::isfinite(dc + dn)) return false;
                const double w = dn / (dc + dn);
                fv = w * field[c] + (1.0 - w) * field[other];
            } else {
                if (m.empty_face(f)) continue;
                const uint8_t t = m.bc_of(f, bct);
                if (t == 1) {
                    fv = bcv[m.patch[f]];
                } else if (t == 2) {
                    double dn = 0.0;
                    if (!boundary_normal_distance(m, f, c, dn)) return false;
                    fv = field[c] + bcv[m.patch[f]] * dn;
                }
            }
            if (!std::isfinite(fv)) return false;
            sum = padd(sum, pmul(sf, fv));
        }
        const double v = m.vol[c];
        if (!(v > 0.0) || !std::isfinite(v)) return false;
        grad[c] = sum.x / v;
        grad[nc + c] = sum.y / v;
        grad[2 * nc + c] = sum.z / v;
    }
    return true;
}

// weighted_least_squares_gradient_robust with INVERSE_DISTANCE_SQUARED
// weights and a requested dimension of 2 or 3:
// core/fvm/least_squares_gradient.h:77-205 (one-sided Jacobi on the weighted
// design matrix, rank from singular values, minimum-norm solve). Returns
// false where the reference throws; *certified mirrors
// quality.full_rank && quality.well_conditioned(limit) (gradient_stencil.h:26-45).
bool robust_wls(const P3& centre, double value, const std::vector<P3>& pts, const std::vector<double>& vals,
                int dim, double limit, P3& g_out, bool& certified) {
    constexpr double eps = std::numeric_limits<double>::epsilon();
    certified = false;
    if (!std::isfinite(value) || !std::isfinite(centre.x) || !std::isfinite(centre.y) || !std::isfinite(centre.z))
        return false;
    const size_t ns = pts.size();
    std::vector<double> design(3 * ns), rhs(ns);
    for (size_t i = 0; i < ns; ++i) {
        const P3 d = psub(pts[i], centre);
        const double r2 = pmag2(d);
        const double dv = vals[i] - value;
        if (!(r2 > 0.0) || !std::isfinite(r2) || !std::isfinite(dv)) return false;
        const double factor = 1.0 / std::sqrt(r2);
        if (!std::isfinite(factor)) return false;
        design[3 * i] = factor * d.x;
        design[3 * i + 1] = factor * d.y;
        design[3 * i + 2] = factor * d.z;
        rhs[i] = factor * dv;
    }
    double V[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    for (size_t sweep = 0; sweep < 40; ++sweep) {
        bool changed = false;
        for (size_t p = 0; p < 2; ++p) {
            for (size_t k = p + 1; k < 3; ++k) {
                double alpha = 0.0, beta = 0.0, gamma = 0.0;
                for (size_t i = 0; i < ns; ++i) {
                    const double* row = &design[3 * i];
                    alpha += row[p] * row[p];
                    beta += row[k] * row[k];
                    gamma += row[p] * row[k];
                }
                if (!(alpha > 0.0) || !(beta > 0.0)) continue;
                if (std::abs(gamma) <= 8.0 * eps * std::sqrt(alpha * beta)) continue;
                const double tau = (beta - alpha) / (2.0 * gamma);
                const double t = std::copysign(1.0, tau) / (std::abs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t);
                const double sn = c * t;
                for (size_t i = 0; i < ns; ++i) {
                    double* row = &design[3 * i];
                    const double op = row[p], ok = row[k];
                    row[p] = c * op - sn * ok;
                    row[k] = sn * op + c * ok;
                }
                for (auto& row : V) {
                    const double op = row[p], ok = row[k];
                    row[p] = c * op - sn * ok;
                    row[k] = sn * op + c * ok;
                }
                changed = true;
            }
        }
        if (!changed) break;
    }
    double sigma[3] = {0.0, 0.0, 0.0};
    for (size_t j = 0; j < 3; ++j) {
        for (size_t i = 0; i < ns; ++i) sigma[j] += design[3 * i + j] * design[3 * i + j];
        sigma[j] = std::sqrt(sigma[j]);
    }
    // Descending order; equal values keep their index order (a three-element
    // std::sort is an insertion sort).
    size_t order[3] = {0, 1, 2};
    for (size_t i = 1; i < 3; ++i)
        for (size_t j = i; j > 0 && sigma[order[j]] > sigma[order[j - 1]]; --j) std::swap(order[j], order[j - 1]);
    const double scale = sigma[order[0]];
    const double cutoff = 128.0 * eps * scale;
    int rank = 0;
    for (size_t j : order)
        if (sigma[j] > cutoff) ++rank;
    if (dim == 2 && rank == 3 && sigma[order[2]] <= 1e-10 * sigma[order[0]]) rank = 2;
    const bool full_rank = rank == dim;
    const double min_pivot = full_rank ? sigma[order[dim - 1]] : 0.0;
    const double cond =
        full_rank && min_pivot > 0.0 ? scale / min_pivot : std::numeric_limits<double>::infinity();
    if (rank != dim) return false;
    double g[3] = {0.0, 0.0, 0.0};
    for (int oi = 0; oi < dim; ++oi) {
        const size_t j = order[oi];
        double proj = 0.0;
        for (size_t i = 0; i < ns; ++i) proj += design[3 * i + j] * rhs[i];
        const double coeff = proj / (sigma[j] * sigma[j]);
        for (size_t k = 0; k < 3; ++k) g[k] += V[k][j] * coeff;
    }
    if (!std::isfinite(g[0]) || !std::isfinite(g[1]) || !std::isfinite(g[2])) return false;
    g_out = {g[0], g[1], g[2]};
    certified = full_rank && rank > 0 && std::isfinite(cond) && cond <= limit;
    return true;
}

// compute_boundary_aware_wls_gradient: finite_volume_transport.h:606-715.
// Returns false on the errors the reference throws; *fallback counts cells
// that keep the Green-Gauss value.
bool wls_gradient(const MeshView& m, const uint8_t* bct, const double* bcv, const double* field, double limit,
                  double* grad, size_t* fallback) {
    if (!gg_gradient(m, bct, bcv, field, grad)) return false;
    int dim = 3;
    for (size_t p = 0; p < m.n_patches; ++p)
        if (m.kind[p] == 1) dim = 2;
    size_t fb = 0;
    std::vector<P3> pts;
    std::vector<double> vals;
    for (size_t c = 0; c < m.nc; ++c) {
        pts.clear();
        vals.clear();
        const P3 centre = at3(m.cc, c);
        const double value = field[c];
        for (uint32_t k = m.coff[c]; k < m.coff[c + 1]; ++k) {
            const size_t f = m.cfaces[k];
            const bool own = m.owner[f] == c;
            if (m.neigh[f] >= 0) {
                const size_t other = own ? size_t(m.neigh[f]) : size_t(m.owner[f]);
                pts.push_back(at3(m.cc, other));
                vals.push_back(field[other]);
                continue;
            }
            if (m.empty_face(f)) continue;
            const uint8_t t = m.bc_of(f, bct);
            const P3 sf = own ? at3(m.fs, f) : pmul(at3(m.fs, f), -1.0);
            const double area = pmag(sf);
            if (!(area > 0.0) || !std::isfinite(area)) return false;
            const P3 normal = pmul(sf, 1.0 / area);
            const double nd = pdot(psub(at3(m.fc, f), centre), normal);
            if (!(nd > 0.0) || !std::isfinite(nd)) return false;
            if (t == 1) {
                pts.push_back(at3(m.fc, f));
                vals.push_back(bcv[m.patch[f]]);
            } else {
                pts.push_back(padd(centre, pmul(normal, nd)));
                vals.push_back(value + nd * (t == 2 ? bcv[m.patch[f]] : 0.0));
            }
        }
        P3 g;
        bool cert = false;
        if (!robust_wls(centre, value, pts, vals, dim, limit, g, cert)) cert = false;
        if (!cert) {
            ++fb;
            continue;
        }
        grad[c] = g.x;
        grad[m.nc + c] = g.y;
        grad[2 * m.nc + c] = g.z;
    }
    if (fallback) *fallback = fb;
    return true;
}

// Per-face tangential diffusion flux for one velocity component with cell
// viscosity mu: compute_nonorthogonal_diffusion_face_fluxes,
// finite_volume_transport.h:717-874 (its WLS gradient of the component
// is passed in).
bool nonorth_fluxes(const MeshView& m, const uint8_t* bct, const double* mu, const double* grad,
                    std::vector<double>& flux) {
    const size_t nc = m.nc;
    flux.assign(m.nf, 0.0);
    bool active = false;
    for (size_t c = 0; c < nc; ++c) {
        if (!(mu[c] >= 0.0) || !std::isfinite(mu[c])) return false;
        active = active || mu[c] > 0.0;
    }
    if (!active) return true;
    for (size_t f = 0; f < m.nf; ++f) {
        const size_t o = m.owner[f];
        const int32_t nr = m.neigh[f];
        if (nr < 0) {
            if (m.empty_face(f)) continue;
            if (m.bc_of(f, bct) != 1) continue;
        }
        const P3 sf = at3(m.fs, f);
        const double area = pmag(sf);
        const P3 d = nr >= 0 ? psub(at3(m.cc, size_t(nr)), at3(m.cc, o)) : psub(at3(m.fc, f), at3(m.cc, o));
        const double sd = pdot(sf, d);
        const double d2 = pdot(d, d);
        if (!(area > 0.0) || !(sd > 0.0) || !(d2 > 0.0) || !std::isfinite(area) || !std::isfinite(sd) ||
            !std::isfinite(d2))
            return false;
        const double alpha = area * area / sd;
        const P3 tangential = psub(sf, pmul(d, alpha));
        if (pmag(tangential) <= 1e-12 * area) continue;
        double gamma_face = mu[o];
        P3 gface{grad[o], grad[nc + o], grad[2 * nc + o]};
        if (nr >= 0) {
            const size_t n = size_t(nr);
            const double area_projection = std::abs(pdot(sf, d)) / area;
            const double d_owner = std::abs(pdot(psub(at3(m.fc, f), at3(m.cc, o)), sf)) / area;
            const double d_neigh = std::abs(pdot(psub(at3(m.cc, n), at3(m.fc, f)), sf)) / area;
            if (!(d_owner > 0.0) || !(d_neigh > 0.0) || !std::isfinite(d_owner) || !std::isfinite(d_neigh) ||
                !(area_projection > 0.0))
                return false;
            double conductance = 0.0;
            const double go = mu[o], gn = mu[n];
            if (go > 0.0 && gn > 0.0) conductance = area / (d_owner / go + d_neigh / gn);
            gamma_face = conductance / alpha;
            const P3 line = psub(at3(m.cc, n), at3(m.cc, o));
            const double d2_line = pmag2(line);
            if (!(d2_line > 0.0) || !std::isfinite(d2_line)) return false;
            const double nw = std::clamp(pdot(psub(at3(m.fc, f), at3(m.cc, o)), line) / d2_line, 0.0, 1.0);
            const double w = 1.0 - nw;
            const P3 gn3{grad[n], grad[nc + n], grad[2 * nc + n]};
            gface = padd(pmul(gface, w), pmul(gn3, 1.0 - w));
        }
        const double corr = gamma_face * pdot(gface, tangential);
        if (!std::isfinite(corr)) return false;
        flux[f] = corr;
    }
    return true;
}

struct Equation {
    std::vector<std::map<size_t, double>> rows;
    std::vector<double> rhs, diag;
};

// assemble_scalar_equation restricted to the momentum predictor's call
// (finite_volume_transport.h:882-1310 via assemble_momentum_component,
// steady_incompressible_solver.h:3262-3311): diffusion_coefficient 0 with
// cell viscosity, UPWIND, no face values/conditions, no extra terms, no
// boundary-face viscosity, non-orthogonal correction on; source_explicit =
// body - grad_p component, source_implicit = 0.
bool assemble_component(const MeshView& m, const uint8_t* bct, const double* bcv, const double* mu,
                        const double* phi, const double* grad_p_d, const double* body_d, const double* u_d,
                        bool bounded, Equation& eq) {
    const size_t nc = m.nc, nf = m.nf;
    bool active = false;
    for (size_t c = 0; c < nc; ++c) {
        if (!(mu[c] >= 0.0) || !std::isfinite(mu[c])) return false;
        active = active || mu[c] > 0.0;
    }
    // needs_nonorthogonal_gradient scan (:987-1021).
    bool needs = false;
    if (active) {
        for (size_t f = 0; f < nf; ++f) {
            const int32_t nr = m.neigh[f];
            if (nr < 0) {
                if (m.empty_face(f)) continue;
                if (m.bc_of(f, bct) != 1) continue;
            }
            const size_t o = m.owner[f];
            const P3 sf = at3(m.fs, f);
            const double area = pmag(sf);
            const P3 d = nr >= 0 ? psub(at3(m.cc, size_t(nr)), at3(m.cc, o)) : psub(at3(m.fc, f), at3(m.cc, o));
            const double sd = pdot(sf, d);
            const double d2 = pdot(d, d);
            if (!(area > 0.0) || !(sd > 0.0) || !(d2 > 0.0) || !std::isfinite(area) || !std::isfinite(sd) ||
                !std::isfinite(d2))
                return false;
            const double alpha = area * area / sd;
            const P3 tangential = psub(sf, pmul(d, alpha));
            if (pmag(tangential) > 1e-12 * area) needs = true;
        }
    }
    std::vector<double> nof(nf, 0.0);
    if (needs) {
        std::vector<double> g(3 * nc);
        if (!wls_gradient(m, bct, bcv, u_d, 1.e8, g.data(), nullptr)) return false;
        if (!nonorth_fluxes(m, bct, mu, g.data(), nof)) return false;
    }
    eq.rows.assign(nc, {});
    std::vector<double> rhs(nc, 0.0), diag(nc, 0.0), div_phi(nc, 0.0), deferred(nc, 0.0);
    for (size_t f = 0; f < nf; ++f) {
        const size_t o = m.owner[f];
        const int32_t nr = m.neigh[f];
        const double F = phi[f];
        if (!std::isfinite(F)) return false;
        if (nr >= 0) {
            const size_t n = size_t(nr);
            const double distance = pmag(psub(at3(m.cc, n), at3(m.cc, o)));
            const double area = pmag(at3(m.fs, f));
            if (!(distance > 0.0) || !(area > 0.0)) return false;
            const P3 Sf = at3(m.fs, f), Cf = at3(m.fc, f), Co = at3(m.cc, o), Cn = at3(m.cc, n);
            const double d_owner = std::abs(pdot(psub(Cf, Co), Sf)) / area;
            const double d_neigh = std::abs(pdot(psub(Cn, Cf), Sf)) / area;
            if (!(d_owner > 0.0) || !(d_neigh > 0.0) || !std::isfinite(d_owner) || !std::isfinite(d_neigh))
                return false;
            double D = 0.0;
            const double go = mu[o], gn = mu[n];
            if (go > 0.0 && gn > 0.0) D = area / (d_owner / go + d_neigh / gn);
            const double a_on = D + std::max(F, 0.0);
            const double a_no = D + std::max(-F, 0.0);
            diag[o] += a_on;
            diag[n] += a_no;
            eq.rows[o][n] -= a_no;
            eq.rows[n][o] -= a_on;
            div_phi[o] += F;
            div_phi[n] -= F;
            if (needs) {
                deferred[o] += nof[f];
                deferred[n] -= nof[f];
            }
        } else {
            if (m.empty_face(f)) continue;
            const uint8_t t = m.bc_of(f, bct);
            const double area = pmag(at3(m.fs, f));
            double distance = 0.0;
            if (!boundary_normal_distance(m, f, o, distance)) return false;
            if (!(distance > 0.0) || !(area > 0.0)) return false;
            if (needs) deferred[o] += nof[f];
            if (t == 1) {
                const double gamma_owner = mu[o];
                const double D = gamma_owner * area / distance;
                const double a_owner = D + std::max(F, 0.0);
                const double a_boundary = D + std::max(-F, 0.0);
                diag[o] += a_owner;
                rhs[o] += a_boundary * bcv[m.patch[f]];
                div_phi[o] += F;
            } else {
                if (bounded) {
                    diag[o] += F;
                } else {
                    diag[o] += std::max(F, 0.0);
                    if (F < 0.0) deferred[o] -= F * u_d[o];
                }
                div_phi[o] += F;
            }
        }
    }
    eq.rhs.assign(nc, 0.0);
    eq.diag.assign(nc, 0.0);
    for (size_t c = 0; c < nc; ++c) {
        if (bounded) diag[c] -= div_phi[c];
        diag[c] -= 0.0 * m.vol[c];
        if (!(diag[c] > 0.0) || !std::isfinite(diag[c])) return false;
        const double source = body_d[c] - grad_p_d[c];
        rhs[c] += source * m.vol[c];
        rhs[c] += deferred[c];
        eq.rows[c][c] += diag[c];
        eq.rhs[c] = rhs[c];
        eq.diag[c] = diag[c];
    }
    return true;
}

// relax_momentum_equation: steady_incompressible_solver.h:3728-3770.
void relax(Equation& eq, const double* u_old, double alpha) {
    if (alpha == 1.0) return;
    for (size_t c = 0; c < eq.diag.size(); ++c) {
        const double a = eq.diag[c];
        const double added = a * (1.0 / alpha - 1.0);
        eq.rows[c][c] += added;
        eq.diag[c] += added;
        eq.rhs[c] += added * u_old[c];
    }
}

MeshView view(size_t nf, const uint32_t* owner, const int32_t* neigh, size_t nc, const uint32_t* coff,
              const uint32_t* cfaces, const double* fc, const double* fs, const double* cc, const double* vol,
              const uint32_t* patch, size_t n_patches, const uint8_t* kind) {
    return MeshView{nf, nc, n_patches, owner, neigh, coff, cfaces, fc, fs, cc, vol, patch, kind};
}

}  // namespace

extern "C" {

// compute_boundary_aware_wls_gradient (finite_volume_transport.h:606-715)
// of a scalar cell field; condition_limit as the reference's argument
// (1e8 by default). grad: 3 x n_cells. *fallback_cells: cells left at the
// Green-Gauss gradient.
int oracle_wls_gradient(size_t nf, const uint32_t* owner, const int32_t* neigh, size_t nc, const uint32_t* coff,
                        const uint32_t* cfaces, const double* fc, const double* fs, const double* cc,
                        const double* vol, const uint32_t* patch, size_t n_patches, const uint8_t* kind,
                        const uint8_t* bc_type, const double* bc_value, const double* field, double condition_limit,
                        double* grad, size_t* fallback_cells) {
    const MeshView m = view(nf, owner, neigh, nc, coff, cfaces, fc, fs, cc, vol, patch, n_patches, kind);
    return wls_gradient(m, bc_type, bc_value, field, condition_limit, grad, fallback_cells) ? 0 : 1;
}

// The three momentum components as the SIMPLE predictor assembles and relaxes
// them (steady_incompressible_solver.h:7173-7193 then :7767-7769): each with
// its own CSR (vals 3 x nnz, component-major), rhs and diagonal (3 x n_cells).
// body may be null (zero body force). Sizes only when rowp is null.
int oracle_momentum_assemble(size_t nf, const uint32_t* owner, const int32_t* neigh, size_t nc,
                             const uint32_t* coff, const uint32_t* cfaces, const double* fc, const double* fs,
                             const double* cc, const double* vol, const uint32_t* patch, size_t n_patches,
                             const uint8_t* kind, const uint8_t* vbc_type, const double* vbc_value, const double* mu,
                             const double* phi, const double* grad_p, const double* body, const double* U,
                             int bounded, double alpha, size_t* nnz_out, uint32_t* rowp, uint32_t* cols,
                             double* vals, double* rhs, double* diag) {
    if (!(alpha > 0.0 && alpha <= 1.0)) return 1;
    const MeshView m = view(nf, owner, neigh, nc, coff, cfaces, fc, fs, cc, vol, patch, n_patches, kind);
    std::vector<double> zero(nc, 0.0), bcv(n_patches);
    std::vector<uint8_t> bct(n_patches);
    for (size_t p = 0; p < n_patches; ++p) bct[p] = vbc_type[p] == 1 ? 1 : 0;
    size_t nnz = 0;
    for (int d = 0; d < 3; ++d) {
        for (size_t p = 0; p < n_patches; ++p) bcv[p] = vbc_value[3 * p + d];
        Equation eq;
        if (!assemble_component(m, bct.data(), bcv.data(), mu, phi, grad_p + d * nc, body ? body + d * nc : zero.data(),
                                U + d * nc, bounded != 0, eq))
            return 1;
        relax(eq, U + d * nc, alpha);
        nnz = 0;
        for (const auto& r : eq.rows) nnz += r.size();
        *nnz_out = nnz;
        if (!rowp) continue;
        size_t k = 0;
        for (size_t r = 0; r < nc; ++r) {
            rowp[r] = uint32_t(k);
            for (const auto& kv : eq.rows[r]) {
                cols[k] = uint32_t(kv.first);
                vals[d * nnz + k] = kv.second;
                ++k;
            }
            rhs[d * nc + r] = eq.rhs[r];
            diag[d * nc + r] = eq.diag[r];
        }
        rowp[nc] = uint32_t(k);
    }
    return 0;
}

// rAU and HbyA of the SIMPLE predictor: rAU = 1 / diagonal
// (steady_incompressible_solver.h:7816-7836), HbyA = rAU * (rhs + grad_p V -
// sum of off-diagonal a_k U_k in storage order) (build_hbya :6430-6465).
// vals/rhs/diag as produced by oracle_momentum_assemble; U: the predicted
// velocity. rAU, HbyA: 3 x n_cells.
int oracle_momentum_hbya(size_t nc, const uint32_t* rowp, const uint32_t* cols, const double* vals,
                         const double* rhs, const double* diag, const double* grad_p, const double* vol,
                         const double* U, double* rAU, double* HbyA) {
    const size_t nnz = rowp[nc];
    for (int d = 0; d < 3; ++d)
        for (size_t c = 0; c < nc; ++c) {
            const double a = diag[d * nc + c];
            if (!(a > 0.0) || !std::isfinite(a)) return 1;
            rAU[d * nc + c] = 1.0 / a;
        }
    for (int d = 0; d < 3; ++d)
        for (size_t c = 0; c < nc; ++c) {
            double h = rhs[d * nc + c] + grad_p[d * nc + c] * vol[c];
            for (uint32_t k = rowp[c]; k < rowp[c + 1]; ++k)
                if (cols[k] != c) h -= vals[d * nnz + k] * U[d * nc + cols[k]];
            HbyA[d * nc + c] = rAU[d * nc + c] * h;
            if (!std::isfinite(HbyA[d * nc + c])) return 1;
        }
    return 0;
}

// Velocity correction U = HbyA - rAU grad_p V: steady_incompressible_solver.h:9545-9557.
void oracle_velocity_correct(size_t nc, const double* HbyA, const double* rAU, const double* grad_p,
                             const double* vol, double* U) {
    for (int d = 0; d < 3; ++d)
        for (size_t c = 0; c < nc; ++c)
            U[d * nc + c] = HbyA[d * nc + c] - rAU[d * nc + c] * grad_p[d * nc + c] * vol[c];
}

// make_rhie_chow_mass_flux with the HBYA_PREDICTOR contract, linear weight
// 1/2, Gauss pressure gradient without skew correction
// (steady_incompressible_solver.h:2063-2217; internal correction
// rhie_chow_pressure_flux_internal :1967-2029). grad_p: the pressure's
// gauss_gradient_with_boundary (oracle_gauss_gradient).
int oracle_rhie_chow_flux(size_t nf, const uint32_t* owner, const int32_t* neigh, size_t nc, const double* fc,
                          const double* fs, const double* cc, const double* vol, const uint32_t* patch,
                          size_t n_patches, const uint8_t* kind, const uint8_t* vbc_type, const double* vbc_value,
                          const uint8_t* pbc_type, const double* pbc_value, double rho, const double* HbyA,
                          const double* p, const double* grad_p, const double* rAU, double* phi) {
    for (int d = 0; d < 3; ++d)
        for (size_t c = 0; c < nc; ++c)
            if (!std::isfinite(rAU[d * nc + c]) || rAU[d * nc + c] <= 0.0) return 1;
    // pressure_face_orthogonal_coefficient :1914-1935 and
    // pressure_face_nonorthogonal_flux :1941-1965.
    auto ortho = [](const P3& Sf, const P3& dv, const P3& dAU, double& out) {
        const double d = pmag(dv);
        const double area = pmag(Sf);
        if (!(d > 0.0) || !std::isfinite(d) || !(area > 0.0) || !std::isfinite(area)) return false;
        const P3 e{dv.x / d, dv.y / d, dv.z / d};
        const double oa = pdot(Sf, e);
        if (!(oa > 0.0) || !std::isfinite(oa)) return false;
        const double rfn = dAU.x * e.x * e.x + dAU.y * e.y * e.y + dAU.z * e.z * e.z;
        out = rfn * oa / d;
        return std::isfinite(out);
    };
    auto nonortho = [](const P3& Sf, const P3& dv, const P3& dAU, const P3& g, double& out) {
        const double d = pmag(dv);
        if (!(d > 0.0) || !std::isfinite(d)) return false;
        const P3 e{dv.x / d, dv.y / d, dv.z / d};
        const double oa = pdot(Sf, e);
        const double rfn = dAU.x * e.x * e.x + dAU.y * e.y * e.y + dAU.z * e.z * e.z;
        const double full = dAU.x * g.x * Sf.x + dAU.y * g.y * Sf.y + dAU.z * g.z * Sf.z;
        const double impl = rfn * pdot(g, e) * oa;
        out = full - impl;
        return std::isfinite(out);
    };
    for (size_t f = 0; f < nf; ++f) {
        const size_t o = owner[f];
        P3 Hf{HbyA[o], HbyA[nc + o], HbyA[2 * nc + o]};
        double correction = 0.0;
        const P3 Sf = at3(fs, f);
        if (neigh[f] >= 0) {
            const size_t n = size_t(neigh[f]);
            const P3 Hn{HbyA[n], HbyA[nc + n], HbyA[2 * nc + n]};
            const double wn = 0.5;
            Hf = padd(pmul(Hf, 1.0 - wn), pmul(Hn, wn));
            const double wo = 1.0 - wn;
            const P3 dAU{wo * vol[o] * rAU[o] + wn * vol[n] * rAU[n],
                         wo * vol[o] * rAU[nc + o] + wn * vol[n] * rAU[nc + n],
                         wo * vol[o] * rAU[2 * nc + o] + wn * vol[n] * rAU[2 * nc + n]};
            const P3 gp{wo * grad_p[o] + wn * grad_p[n], wo * grad_p[nc + o] + wn * grad_p[nc + n],
                        wo * grad_p[2 * nc + o] + wn * grad_p[2 * nc + n]};
            const P3 dv = psub(at3(cc, n), at3(cc, o));
            double ic = 0.0, nof = 0.0;
            if (!ortho(Sf, dv, dAU, ic) || !nonortho(Sf, dv, dAU, gp, nof)) return 1;
            correction = rho * (ic * (p[n] - p[o]) + nof);
        } else {
            const uint32_t pp = patch[f];
            if (pp < n_patches && (kind[pp] == 1 || kind[pp] == 2)) {
                phi[f] = 0.0;
                continue;
            }
            if (pp < n_patches) {
                if (vbc_type[pp] == 1) Hf = {vbc_value[3 * pp], vbc_value[3 * pp + 1], vbc_value[3 * pp + 2]};
                if (pbc_type[pp] == 1) {
                    const P3 dv = psub(at3(fc, f), at3(cc, o));
                    const double d = pmag(dv);
                    if (!(d > 0.0) || !std::isfinite(d)) return 1;
                    const P3 gp{grad_p[o], grad_p[nc + o], grad_p[2 * nc + o]};
                    const P3 dAU{vol[o] * rAU[o], vol[o] * rAU[nc + o], vol[o] * rAU[2 * nc + o]};
                    double ic = 0.0, def = 0.0;
                    if (!ortho(Sf, dv, dAU, ic) || !nonortho(Sf, dv, dAU, gp, def)) return 1;
                    correction = rho * (ic * (pbc_value[pp] - p[o]) + def);
                }
            }
        }
        phi[f] = rho * pdot(Hf, Sf) - correction;
    }
    return 0;
}

}  // extern "C"
