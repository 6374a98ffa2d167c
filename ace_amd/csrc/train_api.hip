// train_api.hip -- host side of the batched training step (included by api.hip after the helpers).

// One ADAM step (ADAM::step, src/opt/algo.cpp:280-296, with the bias corrections of its moments) on the
// FP64 master parameters, in the operation order of the host restatement and without contraction, so
// device and host trajectories agree bit for bit; each thread also writes its parameter into the FP32
// matrix-core layout and, for first-layer weights, the BF16 high/low copies (the indices of param_map).
__global__ __launch_bounds__(256) void k_adam(int n, double* __restrict__ p, double* __restrict__ m,
                                              double* __restrict__ v, const double* __restrict__ g,
                                              double* __restrict__ bt, double gamma, double b1, double b2, double eps,
                                              const int* __restrict__ map, float* __restrict__ wpad,
                                              uint16_t* __restrict__ w0bf, int* __restrict__ done) {
#pragma clang fp contract(off)
    __shared__ int s_last;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) {
        const double c1 = 1.0 / (1.0 - bt[0]), c2 = 1.0 / (1.0 - bt[1]);
        const double gq = g[q];
        const double mq = m[q] * b1 + (1.0 - b1) * gq;
        const double vq = v[q] * b2 + (1.0 - b2) * gq * gq;
        const double pq = p[q] - gamma * c1 * mq / (sqrt(c2 * vq) + eps);
        m[q] = mq;
        v[q] = vq;
        p[q] = pq;
        const int* e = map + 5 * (size_t)q;
        const float x = (float)pq;
        wpad[e[0]] = x;
        if (e[1] >= 0) {
            const uint16_t hi = bf16_rne(x), lo = bf16_rne(x - bf16_value(hi));
            w0bf[e[1]] = hi;
            w0bf[e[2]] = lo;
            w0bf[e[3]] = hi;
            w0bf[e[4]] = lo;
        }
    }
    // (beta1^t, beta2^t) -> (beta1^(t+1), beta2^(t+1)) by the last workgroup, after every workgroup read them
    __syncthreads();
    if (threadIdx.x == 0) s_last = (atomicAdd(done, 1) == (int)gridDim.x - 1);
    __syncthreads();
    if (s_last && threadIdx.x == 0) {
        __threadfence();
        bt[0] *= b1;
        bt[1] *= b2;
        atomicExch(done, 0);
    }
}

#define ADAM_B1 0.9
#define ADAM_B2 0.999

namespace {

void train_release(ace_ctx* ctx) {
    TrainState& T = ctx->tr;
    release(T.d_tiles); release(T.d_Eref); release(T.d_Fref); release(T.d_gfac); release(T.d_eps);
    release(T.d_out); release(T.d_grad); release(T.d_gsum); release(T.d_V); release(T.d_part);
    release(T.d_p); release(T.d_m); release(T.d_v); release(T.d_bt); release(T.d_pmap);
    release(T.d_adone);
    for (int q = 0; q < 2; ++q)
        if (ctx->tr.graph[q]) {
            (void)hipGraphExecDestroy(ctx->tr.graph[q]);
            ctx->tr.graph[q] = nullptr;
        }
}

void train_on_upload(ace_ctx* ctx) {
    ctx->tr.have_targets = false;
    ctx->tr.have_forces = false;
    ctx->tr.g_valid = false;
    ++ctx->seq;
    ctx->tr.tiles.clear();
}

// training tiles: 16 atoms of one species (the matrix-core tile height)
int train_tiles(ace_ctx* ctx) {
    TrainState& T = ctx->tr;
    if (!T.tiles.empty()) return 0;
    const int tm = MLP_ROWS;
    T.tm = tm;
    std::vector<int> cnt(ctx->nt, 0);
    for (int t : ctx->type_h) cnt[t]++;
    T.tile_first.assign(ctx->nt + 1, 0);
    int r = 0;
    for (int I = 0; I < ctx->nt; ++I) {
        T.tile_first[I] = (int)T.tiles.size();
        for (int q = 0; q < cnt[I]; q += tm) T.tiles.push_back(make_int4(I, r + q, std::min(tm, cnt[I] - q), 0));
        r += cnt[I];
    }
    T.tile_first[ctx->nt] = (int)T.tiles.size();
    int ldp = 0;
    for (int I = 0; I < ctx->nt; ++I) ldp = std::max(ldp, ctx->sp[I].nparams());
    T.ldpart = ldp;
    if (upload(ctx, T.d_tiles, T.tiles.data(), T.tiles.size())) return -1;
    if (ensure(ctx, T.d_part, (size_t)T.tiles.size() * ldp)) return -1;
    return 0;
}

int train_grad(ace_ctx* ctx, bool wf);


// The device part of a training step (network, forces, loss, tangent, gradient), after the G rows and the
// buffers are in place: a fixed sequence of launches, captured into a HIP graph when the caller asks.
int train_launch(ace_ctx* ctx, const ace_loss_desc& L, bool wf, bool want_grad) {
    TrainState& T = ctx->tr;
    const int S = ctx->S;
    if (run_mlp(ctx, false)) return -1;
    if (wf) {
        if (run_adjoint(ctx)) return -1;
        ++ctx->seq;
        const int b = 1 - ctx->fbuf;   // the accumulators the adjoint added into
        if (timed(ctx, "loss", [&] {
                k_finalize_loss<<<S, LOSS_THREADS, 0, ctx->stream>>>(
                    ctx->d_atom_off.p, ctx->d_natoms.p, ctx->d_Ffix[b].p, ctx->d_Efix[b].p, ctx->d_Wfix[b].p, ctx->d_F.p,
                    ctx->d_E.p, ctx->d_vir.p, T.d_Eref.p, T.d_Fref.p, L, T.d_gfac.p, T.d_eps.p, T.d_out.p);
            }))
            return -1;
    } else if (timed(ctx, "loss", [&] {
                   k_reduce_loss<<<S, LOSS_THREADS, 0, ctx->stream>>>(ctx->d_atom_off.p, ctx->d_natoms.p,
                                                                      ctx->d_e_atom.p, ctx->d_E.p, T.d_Eref.p, L,
                                                                      T.d_gfac.p, T.d_out.p);
               })) {
        return -1;
    }
    if (wf) {
        if (run_basis(ctx, 2, T.d_V.p, T.d_eps.p)) return -1;
    }
    if (want_grad && train_grad(ctx, wf)) return -1;
    return 0;
}

// parameter gradient from the seeds in gfac (dL/dE per structure) and, when wf, the tangent rows V of the
// force seeds eps (dL/dF): matrix-core pass per tile, then the two-stage reduction into d_grad
int train_grad(ace_ctx* ctx, bool wf) {
    TrainState& T = ctx->tr;
    {
        int lds = 0;
        for (int I = 0; I < ctx->nt; ++I) lds = std::max(lds, train_mc_lds_floats(ctx->hdm.sp[I]) * 4);
        if (lds > 160 * 1024) return fail(ctx, "ace_train_step: network too wide for the training kernel's LDS");
        if (lds > 64 * 1024)
            HC(hipFuncSetAttribute((const void*)k_mlp_train_mc, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        if (timed(ctx, "mlp_train", [&] {
                k_mlp_train_mc<<<(int)T.tiles.size(), 64 * TRAIN_WAVES, lds, ctx->stream>>>(
                    ctx->d_model.p, ctx->d_wpad.p, ctx->d_inpw.p, T.d_tiles.p, ctx->ldg, ctx->d_G.p,
                    wf ? T.d_V.p : nullptr, ctx->d_atom_of_row.p, ctx->d_struc_of.p, T.d_gfac.p, T.ldpart, T.d_part.p);
            }))
            return -1;
        GradRed R;
        memset(&R, 0, sizeof(R));
        R.nt = ctx->nt;
        int off = 0, npmax = 1, ngmax = 1;
        long long gofs = 0;
        for (int I = 0; I < ctx->nt; ++I) {
            R.np[I] = ctx->sp[I].nparams();
            R.off[I] = off;
            R.t0[I] = T.tile_first[I];
            R.t1[I] = T.tile_first[I + 1];
            R.gofs[I] = gofs;
            const int ng = (R.t1[I] - R.t0[I] + GRED_GROUP - 1) / GRED_GROUP;
            gofs += (long long)ng * R.np[I];
            off += R.np[I];
            npmax = std::max(npmax, R.np[I]);
            ngmax = std::max(ngmax, ng);
        }
        // two launches: the kernel boundary orders the group sums before the per-parameter sum at no cost, where
        // one launch would need a device-scope fence per workgroup (an L2 write-back)
        const dim3 g1(cdiv(npmax, 256), ngmax, ctx->nt), g2(cdiv(npmax, 256), 1, ctx->nt);
        if (timed(ctx, "grad_reduce", [&] {
                k_grad_reduce1<<<g1, 256, 0, ctx->stream>>>(R, T.ldpart, T.d_part.p, T.d_gsum.p);
            }) ||
            timed(ctx, "grad_reduce", [&] { k_grad_reduce2<<<g2, 256, 0, ctx->stream>>>(R, T.d_gsum.p, T.d_grad.p); }))
            return -1;
    }
    return 0;
}

// grad-reduction layout of the current tiles: sizes of the pass-1 buffer and the gradient
void train_sizes(ace_ctx* ctx, size_t& ntot, long long& gofs) {
    TrainState& T = ctx->tr;
    ntot = 0;
    gofs = 0;
    for (int I = 0; I < ctx->nt; ++I) {
        const int np = ctx->sp[I].nparams();
        ntot += np;
        gofs += (long long)((T.tile_first[I + 1] - T.tile_first[I] + GRED_GROUP - 1) / GRED_GROUP) * np;
    }
}

// one graph per accumulator buffer the step's force launch adds into (the step toggles it once); captured
// again when a buffer was reallocated or a value baked into the launches changed
int train_graph_launch(ace_ctx* ctx, const ace_loss_desc& L, bool wf, bool want_grad) {
    TrainState& T = ctx->tr;
    const int p = ctx->fbuf;
    long long sig[12];
    memcpy(&sig[0], &L.delta, 8);
    memcpy(&sig[1], &L.w_energy, 8);
    memcpy(&sig[2], &L.w_force, 8);
    sig[3] = (long long)ctx->gen; sig[4] = L.loss; sig[5] = L.norm; sig[6] = wf; sig[7] = want_grad;
    sig[8] = ctx->N; sig[9] = ctx->S; sig[10] = (long long)T.tiles.size() * 64 + ctx->slots_cap;
    sig[11] = ctx->basis_f64;
    if (!T.graph[p] || memcmp(sig, T.graph_sig[p], sizeof(sig)) != 0) {
        if (T.graph[p]) {
            HC(hipGraphExecDestroy(T.graph[p]));
            T.graph[p] = nullptr;
        }
        const bool prof = ctx->prof;
        ctx->prof = false;
        HC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        const int rc = train_launch(ctx, L, wf, want_grad);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        ctx->prof = prof;
        ctx->fbuf = p;
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? -1 : fail_hip(ctx, "train: graph capture", e);
        }
        const hipError_t ei = hipGraphInstantiate(&T.graph[p], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            T.graph[p] = nullptr;
            return fail_hip(ctx, "train: graph instantiate", ei);
        }
        memcpy(T.graph_sig[p], sig, sizeof(sig));
    }
    HC(hipGraphLaunch(T.graph[p], ctx->stream));
    ctx->fbuf = 1 - p;   // as the launches inside did
    return 0;
}

}  // namespace

extern "C" {

int ace_set_targets(ace_ctx* ctx, const double* E_ref, const double* F_ref) {
    if (!ctx) return -1;
    if (ctx->S == 0) return fail(ctx, "ace_set_targets: no structures uploaded");
    if (!E_ref) return fail(ctx, "ace_set_targets: E_ref is required");
    TrainState& T = ctx->tr;
    if (upload(ctx, T.d_Eref, E_ref, ctx->S)) return -1;
    T.have_forces = false;
    if (F_ref) {
        if (upload(ctx, T.d_Fref, F_ref, 3 * (size_t)ctx->N)) return -1;
        T.have_forces = true;
    }
    HC(hipStreamSynchronize(ctx->stream));
    T.have_targets = true;
    return 0;
}

int ace_train_optimizer(ace_ctx* ctx, int algo, double gamma, double eps) {
    if (!ctx) return -1;
    TrainState& T = ctx->tr;
    if (sync_host_params(ctx)) return -1;
    if (algo == ACE_OPT_HOST) {
        T.opt_on = false;
        return 0;
    }
    if (algo != ACE_OPT_ADAM) return fail(ctx, "ace_train_optimizer: only ADAM runs on the device");
    if (!(gamma > 0) || !(eps >= 0)) return fail(ctx, "ace_train_optimizer: invalid gamma or eps");
    if (refresh_params(ctx)) return -1;    // device layouts current before the optimiser writes into them
    const std::vector<int>& map = param_map(ctx);
    std::vector<double> p;
    for (int I = 0; I < ctx->nt; ++I) p.insert(p.end(), ctx->sp[I].params.begin(), ctx->sp[I].params.end());
    const int n = (int)p.size();
    const double bt[2] = {ADAM_B1, ADAM_B2};
    if (upload(ctx, T.d_p, p.data(), n) || upload(ctx, T.d_pmap, map.data(), map.size()) ||
        upload(ctx, T.d_bt, bt, 2) || ensure(ctx, T.d_m, n) || ensure(ctx, T.d_v, n) || ensure(ctx, T.d_adone, 1))
        return -1;
    HC(hipMemsetAsync(T.d_m.p, 0, n * sizeof(double), ctx->stream));
    HC(hipMemsetAsync(T.d_v.p, 0, n * sizeof(double), ctx->stream));
    HC(hipMemsetAsync(T.d_adone.p, 0, sizeof(int), ctx->stream));   // the step's completion counter
    HC(hipStreamSynchronize(ctx->stream));
    T.opt_n = n;
    T.opt_gamma = gamma;
    T.opt_eps = eps;
    T.opt_on = true;
    T.opt_host_stale = false;
    T.opt_dev_stale = false;
    return 0;
}

int ace_train_opt_state(ace_ctx* ctx, int set, double* p, double* m, double* v, double* bt) {
    if (!ctx) return -1;
    TrainState& T = ctx->tr;
    if (!T.opt_on) return fail(ctx, "ace_train_opt_state: no device optimiser");
    if (!p || !m || !v || !bt) return fail(ctx, "ace_train_opt_state: null array");
    const size_t n = T.opt_n;
    if (set) {
        if (upload(ctx, T.d_p, p, n) || upload(ctx, T.d_m, m, n) || upload(ctx, T.d_v, v, n) ||
            upload(ctx, T.d_bt, bt, 2))
            return -1;
        HC(hipStreamSynchronize(ctx->stream));
        size_t q = 0;
        for (int I = 0; I < ctx->nt; ++I)
            for (double& x : ctx->sp[I].params) x = p[q++];
        ctx->params_dirty = true;
        ++ctx->seq;
        ctx->nn64_dirty = true;
        T.opt_host_stale = false;
        T.opt_dev_stale = false;
        return 0;
    }
    if (sync_host_params(ctx)) return -1;
    HC(hipMemcpyAsync(p, T.d_p.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC(hipMemcpyAsync(m, T.d_m.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC(hipMemcpyAsync(v, T.d_v.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC(hipMemcpyAsync(bt, T.d_bt.p, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    return 0;
}

int ace_train_graph(ace_ctx* ctx, int on) {
    if (!ctx) return -1;
    ctx->tr.use_graph = on != 0;
    return 0;
}

int ace_train_step(ace_ctx* ctx, const ace_loss_desc* Ld, double* loss, double* grad) {
    if (!ctx) return -1;
    if (!Ld) return fail(ctx, "ace_train_step: null loss description");
    TrainState& T = ctx->tr;
    if (!T.have_targets) return fail(ctx, "ace_train_step: targets not set");
    const bool wf = Ld->w_force != 0.0;
    if (wf && !T.have_forces) return fail(ctx, "ace_train_step: force term requested without reference forces");
    if (Ld->loss < ACE_LOSS_MSE || Ld->loss > ACE_LOSS_ASINH) return fail(ctx, "ace_train_step: invalid loss");
    if (Ld->norm < ACE_NORM_IDENTITY || Ld->norm > ACE_NORM_LOG) return fail(ctx, "Invalid normalization method.");
    if ((Ld->loss == ACE_LOSS_HUBER || Ld->loss == ACE_LOSS_ASINH) && !(Ld->delta > 0))
        return fail(ctx, "ace_train_step: delta must be positive");
    ctx->basis_f64 = ctx->precision == ACE_PREC_FP64_BASIS;
    if (require_nb(ctx)) return -1;
    // a pending asynchronous rebuild is verified before any training kernel or optimiser update runs; one
    // that outgrew its capacity is rebuilt with exact sizes (as ace_energy_forces does)
    if (ctx->nb_pending && nb_verify(ctx)) {
        if (ctx->nb_async_ok) return -1;
        if (ace_build_neighbors(ctx, ctx->nb_skin) || require_nb(ctx)) return -1;
    }
    if (train_tiles(ctx)) return -1;
    const int N = ctx->N, S = ctx->S;
    if (ensure(ctx, T.d_gfac, S) || ensure(ctx, T.d_eps, 3 * (size_t)N) || ensure(ctx, T.d_out, 3 * (size_t)S))
        return -1;
    HC(hipMemsetAsync(ctx->d_scal.p + 3, 0, sizeof(int), ctx->stream));
    // G depends on positions only: computed once per geometry (the reference precomputes it too)
    if (!T.g_valid) {
        if (run_basis(ctx, 0)) return -1;
        T.g_valid = true;
    }
    const ace_loss_desc L = *Ld;
    size_t ntot = 0;
    long long gofs = 0;
    train_sizes(ctx, ntot, gofs);
    const bool want_grad = grad != nullptr || T.opt_on;
    if ((wf && ensure(ctx, T.d_V, (size_t)N * ctx->ldg)) ||
        (want_grad && (ensure(ctx, T.d_grad, ntot) || ensure(ctx, T.d_gsum, (size_t)std::max(gofs, 1LL)))))
        return -1;
    if (T.opt_on && T.opt_dev_stale) {
        // parameters set on the host since the device optimiser last ran: they become the master copy
        std::vector<double> p;
        for (int I = 0; I < ctx->nt; ++I) p.insert(p.end(), ctx->sp[I].params.begin(), ctx->sp[I].params.end());
        if (upload(ctx, T.d_p, p.data(), p.size())) return -1;
        HC(hipStreamSynchronize(ctx->stream));
        T.opt_dev_stale = false;
    }
    if (T.use_graph) {
        if (train_graph_launch(ctx, L, wf, want_grad)) return -1;
    } else if (train_launch(ctx, L, wf, want_grad)) {
        return -1;
    }
    if (T.opt_on) {
        const int n = T.opt_n;
        if (timed(ctx, "adam", [&] {
                k_adam<<<cdiv(n, 256), 256, 0, ctx->stream>>>(n, T.d_p.p, T.d_m.p, T.d_v.p, T.d_grad.p, T.d_bt.p,
                                                              T.opt_gamma, ADAM_B1, ADAM_B2, T.opt_eps, T.d_pmap.p,
                                                              ctx->d_wpad.p, ctx->d_w0bf.p, T.d_adone.p);
            }))
            return -1;
        T.opt_host_stale = true;
        ctx->nn64_dirty = true;
    }
    if (grad) HC(hipMemcpyAsync(grad, T.d_grad.p, ntot * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    std::vector<double> out(3 * (size_t)S);
    HC(hipMemcpyAsync(out.data(), T.d_out.p, 3 * (size_t)S * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    if (check_err_flag(ctx, "ace_train_step")) return -1;
    if (ctx->nb_pending && nb_verify(ctx)) return -1;
    double le = 0, lf = 0, rm = 0;
    for (int s = 0; s < S; ++s) {
        le += out[3 * s];
        lf += out[3 * s + 1];
        rm += out[3 * s + 2];
    }
    if (loss) {
        loss[0] = le + lf;
        loss[1] = le;
        loss[2] = lf;
        loss[3] = rm;
    }
    return flush_prof(ctx);
}

// Autograd-style pair.  The forward is the split energy/force pass (network-input rows G kept in HBM, network,
// adjoint, fixed-point sums); the backward seeds the training kernels with the caller's dL/dE and dL/dF
// instead of a loss's.
int ace_forward(ace_ctx* ctx, double* E_struc, double* F, double* virial) {
    if (!ctx) return -1;
    TrainState& T = ctx->tr;
    T.fwd_ok = false;
    ctx->basis_f64 = ctx->precision == ACE_PREC_FP64_BASIS;
    if (require_nb(ctx)) return -1;
    if (ctx->nb_pending && nb_verify(ctx)) {
        if (ctx->nb_async_ok) return -1;
        if (ace_build_neighbors(ctx, ctx->nb_skin) || require_nb(ctx)) return -1;
    }
    if (ctx->S == 0) return fail(ctx, "ace_forward: no structures uploaded");
    HC(hipMemsetAsync(ctx->d_scal.p + 3, 0, sizeof(int), ctx->stream));
    if (!T.g_valid) {
        if (run_basis(ctx, 0)) return -1;
        T.g_valid = true;
    }
    if (run_mlp(ctx, false) || run_adjoint(ctx) || run_finalize(ctx)) return -1;
    if (E_struc) HC(hipMemcpyAsync(E_struc, ctx->d_E.p, ctx->S * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (F) HC(hipMemcpyAsync(F, ctx->d_F.p, 3 * (size_t)ctx->N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (virial) HC(hipMemcpyAsync(virial, ctx->d_vir.p, 9 * ctx->S * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    if (check_err_flag(ctx, "ace_forward")) return -1;
    T.fwd_ok = true;
    T.fwd_seq = ctx->seq;
    return flush_prof(ctx);
}

int ace_backward(ace_ctx* ctx, const double* dL_dE, const double* dL_dF, double* dL_dparams, double* dL_dpos) {
    if (!ctx) return -1;
    TrainState& T = ctx->tr;
    if (!T.fwd_ok || T.fwd_seq != ctx->seq)
        return fail(ctx, "ace_backward: no ace_forward at the current positions, parameters, scaling and precision "
                         "(or another device pass ran since)");
    if (dL_dpos && dL_dF)
        return fail(ctx, "ace_backward: dL/dpos through the forces (a Hessian-vector product) is not computed; "
                         "pass dL_dF = NULL for dL/dpos of the energy term");
    const int N = ctx->N, S = ctx->S;
    if (dL_dparams) {
        if (train_tiles(ctx)) return -1;
        size_t ntot = 0;
        long long gofs = 0;
        train_sizes(ctx, ntot, gofs);
        const bool wf = dL_dF != nullptr;
        if (ensure(ctx, T.d_gfac, S) || ensure(ctx, T.d_eps, 3 * (size_t)N) || ensure(ctx, T.d_grad, ntot) ||
            ensure(ctx, T.d_gsum, (size_t)std::max(gofs, 1LL)) || (wf && ensure(ctx, T.d_V, (size_t)N * ctx->ldg)))
            return -1;
        if (dL_dE) {
            HC(hipMemcpyAsync(T.d_gfac.p, dL_dE, S * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        } else {
            HC(hipMemsetAsync(T.d_gfac.p, 0, S * sizeof(double), ctx->stream));
        }
        if (wf) {
            HC(hipMemcpyAsync(T.d_eps.p, dL_dF, 3 * (size_t)N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            if (run_basis(ctx, 2, T.d_V.p, T.d_eps.p)) return -1;
        }
        if (train_grad(ctx, wf)) return -1;
        HC(hipMemcpyAsync(dL_dparams, T.d_grad.p, ntot * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    }
    std::vector<double> F;
    if (dL_dpos) {
        F.resize(3 * (size_t)N);
        HC(hipMemcpyAsync(F.data(), ctx->d_F.p, 3 * (size_t)N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    }
    HC(hipStreamSynchronize(ctx->stream));
    if (check_err_flag(ctx, "ace_backward")) return -1;
    if (dL_dpos) {
        // d/dpos sum_s dL/dE_s E_s = -dL/dE_s F_a for the atoms a of structure s
        for (int st = 0; st < S; ++st)
            for (int a = ctx->atom_off[st]; a < ctx->atom_off[st] + ctx->natoms[st]; ++a)
                for (int d = 0; d < 3; ++d) dL_dpos[3 * a + d] = dL_dE ? -dL_dE[st] * F[3 * a + d] : 0.0;
    }
    T.fwd_seq = ctx->seq;   // the forward stays valid for further backward calls
    return flush_prof(ctx);
}

}  // extern "C"
