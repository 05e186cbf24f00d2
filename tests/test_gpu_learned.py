"""The learned warm start at the parametric-family shape (config 4): a
20,000 x 50,000 LP family c = c0 + P theta, b+- = b0 +- (0.5 + 0.1 |R theta|),
theta in R^16, and the 16 -> 512 -> 512 -> 70,000 MLP (BF16 MFMA, FP32
accumulation) trained on the GPU on solutions the plain PDHG produced.

Checked: the training loss falls; on held-out instances both the cold start
and the learned warm start (with the cold start as its fallback,
sk_gpu_pdlp_session_solve_fallback) reach 1e-6 in the oracle's FP64
relative KKT and agree in objective; a useless warm start triggers the
fallback and still converges.  The gain (cold / warm iterations and wall
time) is measured and printed; it is not asserted, since with the reduced
training set used here (a test must stay short) the network is far from
the accuracy a warm start needs."""
import ctypes as C

import numpy as np
import pytest

import sankhya_gpu as sg
import sk_oracle as so

pytestmark = pytest.mark.gpu

M, N, DIM, SEED = 20000, 50000, 16, 20260829


def _family():
    w = sg.load_workloads()
    model = sg.Model()
    P = np.zeros(N * DIM); R = np.zeros(M * DIM); c0 = np.zeros(N); b0 = np.zeros(M)
    assert w.skw_parametric_family(C.byref(model.m), M, N, 6, DIM, SEED, sg.dptr(P), sg.dptr(R), sg.dptr(c0),
                                   sg.dptr(b0)) == 0
    model._generated = True
    return model, P.reshape(N, DIM), R.reshape(M, DIM), c0, b0


class _Session:
    def __init__(self, model, **kw):
        self.lib = sg.load_gpu()
        p, i, v = model.csc()
        self.rp, self.cj, self.v = sg.csc_to_csr(model.nrow, model.ncol, p, i, v)
        self.l, self.u = model.vec("clow"), model.vec("cupp")
        P = sg.default_pdlp_params(**kw)
        self.h = C.c_void_p()
        vec = [model.vec(k) for k in ("c", "rlow", "rupp", "clow", "cupp")]
        assert self.lib.sk_gpu_pdlp_session_create(model.nrow, model.ncol, int(self.rp[-1]), sg.iptr(self.rp),
                                                   sg.iptr(self.cj), sg.dptr(self.v), *[sg.dptr(a) for a in vec],
                                                   C.byref(P), C.byref(self.h)) == 0, sg.last_error()

    def update(self, c, rl, ru):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (c, rl, ru, self.l, self.u)]
        assert self.lib.sk_gpu_pdlp_session_update(self.h, *[sg.dptr(a) for a in arrs]) == 0

    def solve(self, x0=None, y0=None, warm_budget=0):
        x, y, st = np.zeros(N), np.zeros(M), sg.PdlpStats()
        used = C.c_int(0)
        x0 = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        y0 = None if y0 is None else np.ascontiguousarray(y0, np.float64)
        if warm_budget:
            rc = self.lib.sk_gpu_pdlp_session_solve_fallback(self.h, sg.dptr(x0), sg.dptr(y0), warm_budget,
                                                             sg.dptr(x), sg.dptr(y), C.byref(st), C.byref(used))
        else:
            rc = self.lib.sk_gpu_pdlp_session_solve(self.h, sg.dptr(x0) if x0 is not None else None,
                                                    sg.dptr(y0) if y0 is not None else None, sg.dptr(x),
                                                    sg.dptr(y), C.byref(st))
        assert rc == 0, sg.last_error()
        return x, y, st, used.value

    def close(self):
        self.lib.sk_gpu_pdlp_session_destroy(self.h)


def _kkt(model, c, rl, ru, x, y):
    m2 = sg.Model.from_arrays(M, N, *model.csc(), c, rl, ru, model.vec("clow"), model.vec("cupp"))
    return so.relative_kkt(m2, x, y)


def test_config4_learned_warm_start_with_cold_fallback():
    model, P, R, c0, b0 = _family()
    rng = np.random.default_rng(5)
    train, test = rng.standard_normal((160, DIM)), rng.standard_normal((3, DIM))

    def instance(theta):
        half = 0.5 + 0.1 * np.abs(R @ theta)
        return c0 + P @ theta, b0 - half, b0 + half

    S = _Session(model, precision=sg.PREC_MIXED, use_graph=1, eps_optimal=1e-6, time_limit_seconds=30.0)
    Y = np.zeros((len(train), N + M), np.float32)
    xp = yp = None
    train_its = []
    for k, th in enumerate(train):
        S.update(*instance(th))
        x, y, st, _ = S.solve(xp, yp)       # chained: the previous solution warm-starts the next
        assert st.converged
        Y[k, :N], Y[k, N:] = x, y
        xp, yp = x, y
        train_its.append(st.iterations)
    # the learned start may use twice the training solves' median iteration
    # count before the cold start takes over (known before any test instance)
    warm_budget = int(2 * np.median(train_its))
    X = train.astype(np.float32)
    net = sg.Mlp(DIM, 512, N + M, seed=SEED)
    assert net.n_params == 16 * 512 + 512 + 512 * 512 + 512 + 512 * (N + M) + (N + M)
    ysd = Y.std(0)
    net.set_scaling(X.mean(0), X.std(0), Y.mean(0), np.where(ysd > 1e-6, ysd, 1.0))
    hist = net.train(X, Y, epochs=300, batch=32, lr=1e-3)
    assert np.all(np.isfinite(hist)) and hist[-1] < 0.5 * hist[0], (hist[0], hist[-1])
    pred = net.predict(test.astype(np.float32)).astype(np.float64)
    l, u = model.vec("clow"), model.vec("cupp")
    gains = []
    for k, th in enumerate(test):
        c, rl, ru = instance(th)
        S.update(c, rl, ru)
        xc, yc, stc, _ = S.solve()
        assert stc.converged
        xw, yw, stw, used = S.solve(np.clip(pred[k, :N], l, u), pred[k, N:], warm_budget=warm_budget)
        assert stw.converged
        for xx, yy in ((xc, yc), (xw, yw)):
            kkt = _kkt(model, c, rl, ru, xx, yy)
            assert max(kkt[:3]) <= 1.05e-6, kkt
        oc, ow = float(c @ xc), float(c @ xw)
        assert abs(oc - ow) <= 1e-5 * (1 + abs(oc)), (oc, ow)
        err = np.linalg.norm(pred[k] - np.concatenate([xc, yc])) / np.linalg.norm(np.concatenate([xc, yc]))
        gains.append((stc.iterations, stw.iterations, used, stc.seconds, stw.seconds, err))
    print("config-4 learned warm start (cold it, warm it, warm used, cold s, warm s, prediction error):", gains)
    # a useless start (far outside the solution set) gives up after the warm
    # budget and the cold start solves the instance
    c, rl, ru = instance(test[0])
    S.update(c, rl, ru)
    x, y, st, used = S.solve(np.full(N, 10.0), np.full(M, 1e4), warm_budget=640)
    assert used == 0 and st.converged and st.iterations > 640
    kkt = _kkt(model, c, rl, ru, x, y)
    assert max(kkt[:3]) <= 1.05e-6, kkt
    S.close()
