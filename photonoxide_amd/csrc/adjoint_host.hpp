// The adjoint's host side (pxfdtd_adjoint_*): the forward run with checkpoints, the objective,
// and the reverse sweep with recomputation. Included by resident.hip after adjoint.hpp.
//
// Forward: from step 0 the simulation runs `steps` steps on its device; before each block of K
// steps the state (both fields' current copy and psi) is copied to a checkpoint in HBM. The
// objective is J = |S|^2, S = sum over the monitor's values of m(r) times the FP64 transform sum
// of E's component c at frequency f.
// Backward: dJ/dE after step n at a monitor value is 2 Re(conj(S) m(r) p_n), p_n = e^(i w t) dt
// with t = (n + 1) dt, the phase the transform added there. Block by block from the last, the
// state is restored from its checkpoint and the block's steps recomputed (recording nothing but
// the term cb multiplies at each design value), then the adjoint steps run backwards through
// the block: seed, gradient, then R1a with R1b in one kernel and R2a with R2b in another
// (adjoint.hpp), psibar kept in two buffers that swap each step. dJ/dcb is summed in FP64.
#pragma once

namespace pxfdtd_host {
void sim_set_steps(pxfdtd_simulation *s, int64_t steps);
double sim_dft_frequency(const pxfdtd_simulation *s, int monitor, int f);
int sim_dft_box(const pxfdtd_simulation *s, int monitor, int lo[3], int hi[3]);
bool sim_has_media(const pxfdtd_simulation *s);
bool sim_has_planes(const pxfdtd_simulation *s);
bool sim_e_source_at(const pxfdtd_simulation *s, int comp, long long r);
const double *sim_ca(const pxfdtd_simulation *s);
const double *sim_cb(const pxfdtd_simulation *s);
}  // namespace pxfdtd_host

struct pxfdtd_adjoint {
    pxfdtd_simulation *sim = nullptr;
    pxfdtd_resident *r = nullptr;
    int monitor = 0, component = 0, frequency = 0, K = 64;
    std::vector<double> weights;   // complex m, [k][j][i]
    std::vector<int64_t> mon_index;  // the monitor's values, pitched, within E's component
    std::vector<int64_t> design_idx; // pitched index + component * nval
    std::vector<int8_t> design_comp;
    std::vector<int64_t> design_cell;  // pitched index within a component
    int64_t steps = 0;
    bool forward_done = false;
    double S_re = 0, S_im = 0;
    std::vector<Buf> ckpt;
    size_t ckpt_bytes = 0;
    Buf ca, cb, eb, hb, pb, term, grad, midx, mw, didx, dcomp, dcell;  // pb: two psibar buffers
    int64_t recomputed = 0;
};

template <typename T>
static AdjArgs<T> adj_args(pxfdtd_adjoint *a) {
    pxfdtd_resident *r = a->r;
    AdjArgs<T> A;
    memset(&A, 0, sizeof A);
    AdjGrid &g = A.g;
    g.nx = r->nx; g.ny = r->ny; g.nz = r->nz; g.pitch = r->pitch;
    g.plane = r->plane; g.nval = r->nval;
    for (int w = 0; w < 3; w++) {
        g.per[w] = r->periodic[w];
        g.low[w] = r->low[w];
        g.high[w] = r->high[w];
        g.fe_off[w] = r->fe_off[w];
        g.fh_off[w] = r->fh_off[w];
        A.step[w] = sizeof(T) == 4 ? (T)(1.0 / r->step[w]) : (T)r->step[w];  // see per_step
    }
    memcpy(g.psi_off, r->psi_off, sizeof g.psi_off);
    memcpy(g.cpml_off, r->cpml_off, sizeof g.cpml_off);
    A.dt = (T)r->dt;
    A.coef = r->coef.as<T>();
    A.ca = a->ca.as<T>();
    A.cb = a->cb.as<T>();
    A.eb = a->eb.as<T>();
    A.hb = a->hb.as<T>();
    return A;
}

static unsigned adj_blocks(long long n) { return (unsigned)((n + 255) / 256); }

// adj_h and adj_e's grid: (i, j) in blocks of ADJ_BX x ADJ_BY, k one block row each
static dim3 adj_cell_grid(const pxfdtd_resident *r) {
    return dim3((unsigned)((r->nx + ADJ_BX - 1) / ADJ_BX), (unsigned)((r->ny + ADJ_BY - 1) / ADJ_BY), (unsigned)r->nz);
}

extern "C" int pxfdtd_adjoint_create(pxfdtd_simulation *sim, const pxfdtd_adjoint_spec *spec,
                                     pxfdtd_adjoint **out) {
    pxfdtd_resident *r = pxfdtd_simulation_resident(sim);
    if (!r) return px_fail(PXFDTD_ERR_INVALID, "set the simulation's device before making its adjoint");
    if (pxfdtd_host::sim_steps(sim) != 0)
        return px_fail(PXFDTD_ERR_INVALID, "the adjoint starts from a simulation that has not run");
    if (pxfdtd_host::sim_has_media(sim))
        return px_fail(PXFDTD_ERR_UNSUPPORTED, "the adjoint does not step dispersive media");
    if (pxfdtd_host::sim_has_planes(sim))
        return px_fail(PXFDTD_ERR_UNSUPPORTED, "the adjoint does not checkpoint plane waves' auxiliary grids");
    if (spec->checkpoint < 1) return px_fail(PXFDTD_ERR_INVALID, "checkpoints every K >= 1 steps");
    if (r->nz > 65535 || (r->ny + ADJ_BY - 1) / ADJ_BY > 65535)
        return px_fail(PXFDTD_ERR_UNSUPPORTED, "the adjoint steps at most 65535 z planes and %d y rows", 65535 * ADJ_BY);
    if (spec->component < -1 || spec->component > 2) return px_fail(PXFDTD_ERR_INVALID, "no E component %d", spec->component);
    if (pxfdtd_host::sim_dft_frequency(sim, spec->monitor, spec->frequency) <= 0)
        return px_fail(PXFDTD_ERR_INVALID, "no frequency %d of transform monitor %d", spec->frequency, spec->monitor);
    HIPCHK(hipSetDevice(r->dev->ordinal));
    auto a = std::make_unique<pxfdtd_adjoint>();
    a->sim = sim;
    a->r = r;
    a->monitor = spec->monitor;
    a->component = spec->component;
    a->frequency = spec->frequency;
    a->K = spec->checkpoint;
    int lo[3], hi[3];
    TRY(pxfdtd_host::sim_dft_box(sim, spec->monitor, lo, hi));
    const long long vol = (long long)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    // one component's weights [k][j][i], or (component -1) all three's [c][k][j][i]
    const int c0 = spec->component < 0 ? 0 : spec->component, c1 = spec->component < 0 ? 2 : spec->component;
    a->weights.assign(spec->weights, spec->weights + 2 * vol * (c1 - c0 + 1));
    for (int c = c0; c <= c1; c++)
        for (int k = lo[2]; k < hi[2]; k++)
            for (int j = lo[1]; j < hi[1]; j++)
                for (int i = lo[0]; i < hi[0]; i++)
                    a->mon_index.push_back(c * r->nval + (long long)k * r->plane + (long long)j * r->pitch + i);
    const long long N = r->ncells;
    const double *ca = pxfdtd_host::sim_ca(sim), *cb = pxfdtd_host::sim_cb(sim);
    for (int64_t d = 0; d < spec->ndesign; d++) {
        const int c = spec->design_component[d];
        const long long v = spec->design_index[d];
        if (c < 0 || c > 2 || v < 0 || v >= N) return px_fail(PXFDTD_ERR_INVALID, "design value %lld is off the grid", (long long)d);
        const int i = (int)(v % r->nx), j = (int)((v / r->nx) % r->ny), k = (int)(v / ((long long)r->nx * r->ny));
        const int m3[3] = {i, j, k};
        const int n3[3] = {r->nx, r->ny, r->nz};
        for (int w = 0; w < 3; w++)
            if ((r->low[w] > 0 && m3[w] < r->low[w]) || (r->high[w] > 0 && m3[w] >= n3[w] - r->high[w]))
                return px_fail(PXFDTD_ERR_UNSUPPORTED, "design value %lld lies in a CPML slab", (long long)d);
        if (ca[c * N + v] != 1.0 || cb[c * N + v] == 0.0)
            return px_fail(PXFDTD_ERR_UNSUPPORTED, "design value %lld is lossy, a wall or a conductor", (long long)d);
        if (pxfdtd_host::sim_e_source_at(sim, c, v))
            return px_fail(PXFDTD_ERR_UNSUPPORTED, "design value %lld carries an E source", (long long)d);
        const long long cell = (long long)k * r->plane + (long long)j * r->pitch + i;
        a->design_cell.push_back(cell);
        a->design_comp.push_back((int8_t)c);
        a->design_idx.push_back(c * r->nval + cell);
    }
    *out = a.release();
    return PXFDTD_OK;
}

extern "C" void pxfdtd_adjoint_destroy(pxfdtd_adjoint *a) { delete a; }

template <typename T>
static int adj_save(pxfdtd_adjoint *a, size_t slot) {
    pxfdtd_resident *r = a->r;
    const size_t fb = 6 * (size_t)r->nval * sizeof(T), pbytes = (size_t)r->psi_len * sizeof(T);
    if (a->ckpt.size() <= slot) {
        std::vector<Buf> grown(slot + 1);
        for (size_t q = 0; q < a->ckpt.size(); q++) grown[q].swap(a->ckpt[q]);
        a->ckpt.swap(grown);
    }
    TRY(a->ckpt[slot].reserve(fb + pbytes));
    HIPCHK(hipMemcpyAsync(a->ckpt[slot].p, r->fields[r->p].p, fb, hipMemcpyDeviceToDevice, r->stream));
    if (pbytes)
        HIPCHK(hipMemcpyAsync((char *)a->ckpt[slot].p + fb, r->psi[r->p].p, pbytes, hipMemcpyDeviceToDevice, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    return PXFDTD_OK;
}

template <typename T>
static int adj_restore(pxfdtd_adjoint *a, size_t slot) {
    pxfdtd_resident *r = a->r;
    const size_t fb = 6 * (size_t)r->nval * sizeof(T), pbytes = (size_t)r->psi_len * sizeof(T);
    HIPCHK(hipMemcpyAsync(r->fields[r->p].p, a->ckpt[slot].p, fb, hipMemcpyDeviceToDevice, r->stream));
    if (pbytes)
        HIPCHK(hipMemcpyAsync(r->psi[r->p].p, (char *)a->ckpt[slot].p + fb, pbytes, hipMemcpyDeviceToDevice, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    return PXFDTD_OK;
}

template <typename T>
static int adj_forward_t(pxfdtd_adjoint *a, int64_t steps) {
    const int K = a->K;
    int64_t slot = 0;
    for (int64_t done = 0; done < steps; slot++) {
        TRY(adj_save<T>(a, (size_t)slot));
        const int64_t m = std::min<int64_t>(K, steps - done);
        TRY(pxfdtd_simulation_run(a->sim, m));
        done += m;
    }
    TRY(adj_save<T>(a, (size_t)slot));  // the end state, put back after the backward sweep
    a->ckpt_bytes = 0;
    for (auto &b : a->ckpt) a->ckpt_bytes += b.bytes;
    return PXFDTD_OK;
}

extern "C" int pxfdtd_adjoint_forward(pxfdtd_adjoint *a, int64_t steps, double *objective) {
    if (steps < 1) return px_fail(PXFDTD_ERR_INVALID, "the forward run needs steps");
    if (a->forward_done) return px_fail(PXFDTD_ERR_INVALID, "the forward run has been done");
    if (pxfdtd_host::sim_has_planes(a->sim))
        return px_fail(PXFDTD_ERR_UNSUPPORTED, "the adjoint does not checkpoint plane waves' auxiliary grids");
    HIPCHK(hipSetDevice(a->r->dev->ordinal));
    TRY(a->r->prec == PXFDTD_F64 ? adj_forward_t<double>(a, steps) : adj_forward_t<float>(a, steps));
    a->steps = steps;
    a->forward_done = true;
    // S = sum m(r) * transform of E_c at frequency f, over one component or all three
    const int c0 = a->component < 0 ? 0 : a->component, c1 = a->component < 0 ? 2 : a->component;
    const long long vol = (long long)a->mon_index.size() / (c1 - c0 + 1);
    double re = 0, im = 0;
    for (int c = c0; c <= c1; c++) {
        const int64_t len = pxfdtd_simulation_dft_len(a->sim, a->monitor, c);
        std::vector<double> sums((size_t)len);
        TRY(pxfdtd_simulation_dft(a->sim, a->monitor, c, sums.data()));
        const double *s = sums.data() + 2 * (size_t)a->frequency * vol;
        const double *w = a->weights.data() + 2 * (size_t)(c - c0) * vol;
        for (long long v = 0; v < vol; v++) {
            const double mr = w[2 * v], mi = w[2 * v + 1];
            re += mr * s[2 * v] - mi * s[2 * v + 1];
            im += mr * s[2 * v + 1] + mi * s[2 * v];
        }
    }
    a->S_re = re;
    a->S_im = im;
    *objective = re * re + im * im;
    return PXFDTD_OK;
}

template <typename T>
static int adj_backward_t(pxfdtd_adjoint *a, double *grad_out) {
    pxfdtd_resident *r = a->r;
    pxfdtd_simulation *sim = a->sim;
    hipStream_t st = r->stream;
    const long long nval = r->nval, D = (long long)a->design_idx.size(), M = (long long)a->mon_index.size();
    const int K = a->K;
    // coefficients [3][nval] pitched, the adjoint state, tables
    {
        std::vector<T> ca(3 * nval, T(0)), cb(3 * nval, T(0));
        const double *hca = pxfdtd_host::sim_ca(sim), *hcb = pxfdtd_host::sim_cb(sim);
        for (int c = 0; c < 3; c++) {
            to_pitched<T>(r, hca + c * r->ncells, ca.data() + c * nval);
            to_pitched<T>(r, hcb + c * r->ncells, cb.data() + c * nval);
        }
        TRY(a->ca.alloc(ca.size() * sizeof(T)));
        TRY(a->cb.alloc(cb.size() * sizeof(T)));
        HIPCHK(hipMemcpy(a->ca.p, ca.data(), ca.size() * sizeof(T), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(a->cb.p, cb.data(), cb.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    TRY(a->eb.alloc(3 * nval * sizeof(T)));
    TRY(a->hb.alloc(3 * nval * sizeof(T)));
    const long long plen = std::max<long long>(r->psi_len, 1);
    TRY(a->pb.alloc(2 * plen * sizeof(T)));
    HIPCHK(hipMemsetAsync(a->eb.p, 0, a->eb.bytes, st));
    HIPCHK(hipMemsetAsync(a->hb.p, 0, a->hb.bytes, st));
    HIPCHK(hipMemsetAsync(a->pb.p, 0, a->pb.bytes, st));
    // the seed's weights: w = conj(S) m
    std::vector<double> w(2 * M);
    for (long long v = 0; v < M; v++) {
        const double mr = a->weights[2 * v], mi = a->weights[2 * v + 1];
        w[2 * v] = a->S_re * mr + a->S_im * mi;
        w[2 * v + 1] = a->S_re * mi - a->S_im * mr;
    }
    TRY(upload_vec(a->midx, a->mon_index, st));
    TRY(upload_vec(a->mw, w, st));
    TRY(upload_vec(a->didx, a->design_idx, st));
    TRY(upload_vec(a->dcomp, a->design_comp, st));
    TRY(upload_vec(a->dcell, a->design_cell, st));
    TRY(a->term.alloc(std::max<long long>((long long)K * D, 1) * sizeof(T)));
    TRY(a->grad.alloc(std::max<long long>(D, 1) * sizeof(double)));
    HIPCHK(hipMemsetAsync(a->grad.p, 0, a->grad.bytes, st));
    HIPCHK(hipStreamSynchronize(st));
    AdjArgs<T> A = adj_args<T>(a);
    const double dt = r->dt;
    const double omega = 2.0 * 3.141592653589793 * pxfdtd_host::sim_dft_frequency(sim, a->monitor, a->frequency);
    const int slots = std::max(1, (int)r->slots);
    const int em = std::max(1, pxfdtd_host::sim_emitters(sim));
    std::vector<double> waves((size_t)em * 2 * K), phases((size_t)slots * 2 * K);
    const int64_t blocks = (a->steps + K - 1) / K;
    const dim3 cgrid = adj_cell_grid(r), cblock(ADJ_BX, ADJ_BY);
    int cur = 0;  // which psibar buffer holds the state
    r->quiet = true;
    int rc = PXFDTD_OK;
    for (int64_t b = blocks - 1; b >= 0 && rc == PXFDTD_OK; b--) {
        const int64_t n0 = b * K, m = std::min<int64_t>(K, a->steps - n0);
        rc = adj_restore<T>(a, (size_t)b);
        if (rc) break;
        // recompute the block in one call, recording after each step the term cb multiplies at
        // the design values (from the step's new H)
        pxfdtd_host::sim_tables(sim, n0, (int)m, slots, waves.data(), phases.data());
        if (D)
            r->after_step = [&](hipStream_t on, uint32_t qn) {
                hipLaunchKernelGGL((adj_term<T>), dim3(adj_blocks(D)), dim3(256), 0, on, A.g, A.coef,
                                   r->field<T>(r->p, 3), a->dcell.as<int64_t>(), a->dcomp.as<int8_t>(), D,
                                   a->term.as<T>() + (int64_t)qn * D);
            };
        rc = pxfdtd_resident_step(r, (uint32_t)m, waves.data(), phases.data());
        r->after_step = nullptr;
        a->recomputed += m;
        if (rc) break;
        // the adjoint steps backwards through the block
        for (int64_t qn = m - 1; qn >= 0; qn--) {
            const double t = (double)(n0 + qn + 1) * dt;
            const double pr = std::cos(omega * t) * dt, pi = std::sin(omega * t) * dt;
            if (M)
                hipLaunchKernelGGL((adj_seed<T>), dim3(adj_blocks(M)), dim3(256), 0, st, A.eb, a->midx.as<int64_t>(),
                                   a->mw.as<double>(), M, pr, pi);
            if (D)
                hipLaunchKernelGGL((adj_grad<T>), dim3(adj_blocks(D)), dim3(256), 0, st, A.eb, a->didx.as<int64_t>(),
                                   a->term.as<T>() + qn * D, D, a->grad.as<double>());
            A.pb_in = a->pb.as<T>() + cur * plen;
            A.pb_out = a->pb.as<T>() + (cur ^ 1) * plen;
            hipLaunchKernelGGL((adj_h<T>), cgrid, cblock, 0, st, A);
            hipLaunchKernelGGL((adj_e<T>), cgrid, cblock, 0, st, A);
            cur ^= 1;
        }
        const hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = px_fail(PXFDTD_ERR_DEVICE, "the adjoint steps: %s", hipGetErrorString(e));
    }
    r->quiet = false;
    if (rc == PXFDTD_OK) rc = adj_restore<T>(a, (size_t)blocks);  // the simulation as the forward left it
    pxfdtd_host::sim_set_steps(sim, a->steps);
    TRY(rc);
    HIPCHK(hipMemcpy(grad_out, a->grad.p, (size_t)D * sizeof(double), hipMemcpyDeviceToHost));
    return PXFDTD_OK;
}

extern "C" int pxfdtd_adjoint_backward(pxfdtd_adjoint *a, double *grad_cb) {
    if (!a->forward_done) return px_fail(PXFDTD_ERR_INVALID, "run the forward before the backward");
    HIPCHK(hipSetDevice(a->r->dev->ordinal));
    return a->r->prec == PXFDTD_F64 ? adj_backward_t<double>(a, grad_cb) : adj_backward_t<float>(a, grad_cb);
}

extern "C" int pxfdtd_adjoint_info(const pxfdtd_adjoint *a, int64_t *checkpoints, int64_t *checkpoint_bytes,
                                   int64_t *recomputed_steps) {
    *checkpoints = (int64_t)a->ckpt.size() - (a->ckpt.empty() ? 0 : 1);  // without the end state
    *checkpoint_bytes = (int64_t)a->ckpt_bytes;
    *recomputed_steps = a->recomputed;
    return PXFDTD_OK;
}
