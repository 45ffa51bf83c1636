// RHF nuclear gradient (host side).  The reference's
// HartreeFock/Derivatives.hs only sketches derivative integrals and does
// not compile, so this entry point has no reference counterpart; it
// differentiates the energy the rest of the library computes:
//   E = 1/2 sum D (H + F) + V_nn   (HartreeFock.hs:70-123, variationalE)
//   dE/dR = sum D dH/dR + 1/2 sum Gamma d(ij|kl)/dR - sum W dS/dR + dV_nn/dR
// with W = 1/2 D F D (= 2 sum_occ eps_i C_i C_i^T at convergence).
#include <math.h>

#include <algorithm>
#include <vector>

#include "hfk_host.hpp"

static int grad_upload_maps(hfk_ctx* c)
{
    if (c->d_sh_atom) return HFK_OK;
    const int natom = (int)c->Z.size();
    std::vector<int> sa(c->shells.size());
    for (size_t s = 0; s < c->shells.size(); ++s) {
        sa[s] = c->shells[s].atom;
        if (sa[s] < 0 || sa[s] >= natom)
            return hfk_set_error(HFK_EINVAL, "hfk_gradient: shell %zu has atom %d (natom %d)", s, sa[s], natom);
    }
    for (int a : c->atom_of_bf)
        if (a < 0 || a >= natom) return hfk_set_error(HFK_EINVAL, "hfk_gradient: atom_of_bf out of range");
    HFK_TRY(dev_upload(c, &c->d_sh_atom, sa));
    HFK_TRY(dev_upload(c, &c->d_atom_of_bf, c->atom_of_bf));
    return HFK_OK;
}

// W = 1/2 D F D with F = H + G(D); D, W flat on device.  After a converged
// hfk_scf the context holds a density D0 and its Fock matrix F(D0); since G is
// linear, F(D) = F(D0) + G(D - D0), built with the Schwarz x max|D - D0| screen
// at tau = 1e-13 (each skipped shell quartet changes an element of F by less
// than 4e-13).  Near convergence that costs a small part of a full build.
// The context's gradient workspace (allocated on the first hfk_gradient and
// kept): H, F (flat), Df, Ff, T (n*n each; T also holds reduction partials),
// the per-atom sums.
static int grad_workspace(hfk_ctx* c, int natom)
{
    if (c->d_gwork) return HFK_OK;
    const size_t nf = (size_t)c->n * (c->n + 1) / 2, nn = std::max<size_t>((size_t)c->n * c->n, 1024);
    return dev_alloc(&c->d_gwork, 2 * nf + 3 * nn + 3 * (size_t)natom);
}

static int grad_energy_weighted(hfk_ctx* c, const double* Dflat, double* Wflat)
{
    const int n = c->n;
    const size_t nf = (size_t)n * (n + 1) / 2, nn = std::max<size_t>((size_t)n * n, 1024);
    double* H = c->d_gwork;
    double* F = H + nf;
    double* Df = F + nf;
    double* Ff = Df + nn;
    double* T = Ff + nn;   // also the reduction partials
    int rc = HFK_OK;
    if (c->scf_cached) {
        // Df (n*n >= nf) holds D - D0, T the reduction partials, Ff[0] max|D - D0|
        hipMemcpyAsync(Df, Dflat, sizeof(double) * nf, hipMemcpyDeviceToDevice, c->stream);
        hfk_launch_axpby((long long)nf, -1.0, c->d_scf_D, 1.0, Df, c->stream);
        hfk_launch_reduce((long long)nf, Df, nullptr, 1, T, Ff, c->stream);
        double dmax = 0.0;
        if (hipMemcpyAsync(&dmax, Ff, sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = hfk_set_error(HFK_EHIP, "hfk_gradient: density difference failed");
        if (rc == HFK_OK) {
            if (dmax > 0.0) rc = hfk_ctx_fock_device(c, c->d_scf_F, Df, F, 1e-13 / dmax);
            else if (hipMemcpyAsync(F, c->d_scf_F, sizeof(double) * nf, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
                rc = hfk_set_error(HFK_EHIP, "hfk_gradient: copy failed");
        }
    } else {
        if (rc == HFK_OK) rc = hfk_ctx_hcore_dev(c, H);
        if (rc == HFK_OK) rc = hfk_ctx_fock_device(c, H, Dflat, F);
    }
    if (rc == HFK_OK) {
        hfk_launch_flat_to_full(n, Dflat, Df, c->stream);
        hfk_launch_flat_to_full(n, F, Ff, c->stream);
        hfk_launch_dgemm(n, n, n, 1.0, Ff, n, 0, Df, n, 0, 0.0, T, n, c->stream);    // F D
        hfk_launch_dgemm(n, n, n, 0.5, Df, n, 0, T, n, 0, 0.0, Ff, n, c->stream);   // 1/2 D F D
        hfk_launch_full_to_flat(n, Ff, Wflat, c->stream);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
            rc = hfk_set_error(HFK_EHIP, "hfk_gradient: energy-weighted density failed");
    }
    return rc;
}

extern "C" int hfk_gradient(hfk_ctx* c, const double* D_flat, const double* W_flat, double* g)
{
    if (!c || !D_flat || !g) return hfk_set_error(HFK_EINVAL, "hfk_gradient: null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    const int n = c->n, natom = (int)c->Z.size();
    if (natom == 0 || (int)c->atom_of_bf.size() != n)
        return hfk_set_error(HFK_EINVAL, "hfk_gradient: basis needs natom, Z, R and atom_of_bf");
    if (hfk_grad_lds_bytes(0, 0, 0, natom, 0, 0) > 64 * 1024)
        return hfk_set_error(HFK_EUNSUP, "hfk_gradient: %d atoms exceed the per-workgroup accumulator", natom);
    HFK_TRY(grad_upload_maps(c));
    HFK_TRY(grad_workspace(c, natom));
    const size_t nf = (size_t)n * (n + 1) / 2;
    HFK_HIP(hipMemcpyAsync(c->d_flatA, D_flat, sizeof(double) * nf, hipMemcpyHostToDevice, c->stream));
    if (W_flat) {
        HFK_HIP(hipMemcpyAsync(c->d_flatB, W_flat, sizeof(double) * nf, hipMemcpyHostToDevice, c->stream));
    } else {
        HFK_TRY(grad_energy_weighted(c, c->d_flatA, c->d_flatB));
    }
    double* d_grad = c->d_gwork + 2 * ((size_t)n * (n + 1) / 2) + 3 * std::max<size_t>((size_t)n * n, 1024);
    hfk_launch_fill(d_grad, 0.0, 3LL * natom, c->stream);
    hfk_launch_flat_to_padded(n, c->d_flatA, c->d_D, c->stream);
    QuartetArgs a = hfk_ctx_args(c);
    a.sch_cut = c->opt.schwarz_cut;
    a.D = c->d_D;
    a.grad = d_grad;
    a.sh_atom = c->d_sh_atom;
    a.natom = natom;
    a.grad_lds = 1;
    int rc = HFK_OK;
    // the class launches run one after another on the context stream: each
    // fills the GPU, and at (H2O)_64 launching them on several streams was
    // slower (1.107 s against 1.091 s)
    const size_t np = c->plan.size();
    for (size_t i = 0; i < np; ++i) {
        const PlanEntry& e = c->plan[i];
        hfk_plan_args(c, e, a);
        // the bra shells' density rows in LDS while four waves per SIMD still fit
        a.drows = !hfk_class_grad_coop(e.la, e.lb, e.lc, e.ld) &&
                  hfk_grad_lds_bytes(e.la, e.lb, e.ncols, natom, 1, e.bra_stage) <= hfk_lds_budget(a.block);
        if (!hfk_class_grad_coop(e.la, e.lb, e.lc, e.ld) &&
            hfk_grad_lds_bytes(e.la, e.lb, e.ncols, natom, 0, e.bra_stage) > 64 * 1024) {
            rc = hfk_set_error(HFK_EUNSUP, "hfk_gradient: %d atoms exceed the per-workgroup accumulator", natom);
            break;
        }
        hipStream_t s = c->stream;
        HFK_HIP(hipEventRecord(c->events[4 * np + 2 * i], s));
        // a mixed-precision task list: its FP64 part in FP64, its FP32 part in
        // FP32 (per-lane kernels); the cooperative kernel takes the whole list
        // in FP64
        const bool parts = e.split && !hfk_class_grad_coop(e.la, e.lb, e.lc, e.ld);
        bool ok = true;
        for (int part = 0; part < (parts ? 2 : 1) && ok; ++part) {
            QuartetArgs b = a;
            if (parts) {
                const long long t0 = part == 0 ? 0 : e.ntask64;
                b.tasks = e.d_tasks + t0;
                b.task0 = t0;
                b.ntask = part == 0 ? e.ntask64 : e.ntask - e.ntask64;
            }
            ok = hfk_launch_grad(e.la, e.lb, e.lc, e.ld, c->opt.boys_mode, e.mode, part, b, s) == 0;
        }
        if (!ok) {
            rc = hfk_set_error(HFK_EUNSUP, "no gradient kernel for class (%d%d|%d%d)", e.la, e.lb, e.lc, e.ld);
            break;
        }
        HFK_HIP(hipEventRecord(c->events[4 * np + 2 * i + 1], s));
    }
    if (rc == HFK_OK) {
        hfk_launch_grad_onee(hfk_onee_args(c), c->d_atom_of_bf, c->d_flatA, c->d_flatB, d_grad, c->stream);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(g, d_grad, sizeof(double) * 3 * natom, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = hfk_set_error(HFK_EHIP, "hfk_gradient: kernel or copy-out failed");
    }
    if (rc != HFK_OK) return rc;
    // per-class launch times
    for (size_t i = 0; i < np; ++i) {
        float ms = 0.f;
        HFK_HIP(hipEventElapsedTime(&ms, c->events[4 * np + 2 * i], c->events[4 * np + 2 * i + 1]));
        c->plan[i].grad_ms = ms;
    }
    // nuclear repulsion: d/dR_A sum_{B<C} Z_B Z_C / |R_B - R_C|
    for (int A = 0; A < natom; ++A)
        for (int B = 0; B < natom; ++B) {
            if (A == B) continue;
            double d[3], r2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                d[k] = c->R[3 * A + k] - c->R[3 * B + k];
                r2 += d[k] * d[k];
            }
            const double r = sqrt(r2), f = -c->Z[A] * c->Z[B] / (r2 * r);
            for (int k = 0; k < 3; ++k) g[3 * A + k] += f * d[k];
        }
    return HFK_OK;
}
