"""Closure pipeline: fine runs -> restricted training pairs -> training ->
coarse+closure runs judged against the fine solver (no counterpart in the
reference).

Grids: the fine basin has N_f = 2*N_c cells per side at dx_f and time step
rdt_f; the coarse basin N_c at dx_c = 2*dx_f, rdt_c = 2*rdt_f, so one coarse
step covers two fine steps and both see the same tide at the same time.
Both carry seeded Gaussian seamounts (nemo2d.seamount_grid on the fine grid,
restricted to the coarse one).

Restriction R (fine -> coarse) on the C-grid:
  ssh_c  = mean of the 4 fine T cells (volume-conserving: sum e12t*ssh is
           unchanged);
  u_c    = depth-weighted mean of the 2 fine U faces lying on the coarse U
           face (the coarse face carries the fine faces' volume flux when the
           coarse depth is their mean); v_c likewise;
  sshn_u, sshn_v are re-derived from ssh_c with the serial model's next
  rules (original/nemolite2d.f90:735-765), row 0 of sshn_v staying 0.
Training pairs: from C_t = R(F_t) the coarse model runs K = 10 steps to C';
the target is R(F_{t+K}) - C' and the input is C' plus the depth, one
pixel per T cell carrying the cell's own C-grid points: ssh at the centre,
u on its east face, v on its north face (no averaging, so an increment
lands exactly where it was measured; the west and south outer faces, one a
wall and one the open boundary's Flather face, are never corrected, and the
east and north wall faces stay 0).  Applying the closure adds the predicted
increment every K coarse steps.
"""
import numpy as np

from . import Model, default_grid, initial_state, seamount_grid, shapes

K_STEPS = 10


# ---- grids and restriction ----------------------------------------------------
def face_depths(ht):
    jpj, jpi = ht.shape
    hu = np.empty((jpj, jpi + 1))
    hu[:, 1:jpi] = 0.5 * (ht[:, :-1] + ht[:, 1:])
    hu[:, 0], hu[:, jpi] = ht[:, 0], ht[:, -1]
    hv = np.empty((jpj + 1, jpi))
    hv[1:jpj, :] = 0.5 * (ht[:-1, :] + ht[1:, :])
    hv[0, :], hv[jpj, :] = ht[0, :], ht[-1, :]
    return hu, hv


def coarse_grid(fine, dx_c, dy_c):
    nf = fine["ht"].shape[0]
    nc = nf // 2
    g = default_grid(nc, nc, dx_c, dy_c, 100.0)
    htf = fine["ht"]
    g["ht"] = 0.25 * (htf[0::2, 0::2] + htf[0::2, 1::2] + htf[1::2, 0::2] + htf[1::2, 1::2])
    g["hu"], g["hv"] = face_depths(g["ht"])
    return g


def next_interp(ssh, e12t=None, e12u=None, e12v=None):
    """sshn_u, sshn_v from ssh by the serial model's next rules on the
    reference basin (uniform areas): interior faces average, outer faces copy
    the adjacent cell, V row 0 (open boundary) stays 0."""
    jpj, jpi = ssh.shape
    su = np.empty((jpj, jpi + 1))
    sv = np.zeros((jpj + 1, jpi))
    e = 1.0 if e12t is None else e12t
    su[:, 1:jpi] = 0.5 * (e * ssh[:, :-1] + e * ssh[:, 1:]) / (1.0 if e12u is None else e12u)
    su[:, 0], su[:, jpi] = ssh[:, 0], ssh[:, -1]
    sv[1:jpj, :] = 0.5 * (e * ssh[:-1, :] + e * ssh[1:, :]) / (1.0 if e12v is None else e12v)
    sv[jpj, :] = ssh[-1, :]
    return su, sv


def restrict_state(fine_state, fine_grid, coarse_g):
    f = fine_state
    ssh = 0.25 * (f["sshn"][0::2, 0::2] + f["sshn"][0::2, 1::2] + f["sshn"][1::2, 0::2] +
                  f["sshn"][1::2, 1::2])
    huf, hvf = fine_grid["hu"], fine_grid["hv"]
    u = f["un"][:, 0::2]
    hu = huf[:, 0::2]
    un = (hu[0::2] * u[0::2] + hu[1::2] * u[1::2]) / (hu[0::2] + hu[1::2])
    v = f["vn"][0::2, :]
    hv = hvf[0::2, :]
    vn = (hv[:, 0::2] * v[:, 0::2] + hv[:, 1::2] * v[:, 1::2]) / (hv[:, 0::2] + hv[:, 1::2])
    e = coarse_g["e12t"][0, 0]
    su, sv = next_interp(ssh, e, coarse_g["e12u"][0, 0], coarse_g["e12v"][0, 0])
    return {"sshn": ssh, "sshn_u": su, "sshn_v": sv, "un": np.ascontiguousarray(un),
            "vn": np.ascontiguousarray(vn)}


def t_fields(state):
    """Per T cell: ssh, u on the east face, v on the north face."""
    return np.stack([state["sshn"], state["un"][:, 1:], state["vn"][1:, :]], -1)


def features(state, ht):
    return np.concatenate([t_fields(state), ht[..., None]], -1)


def apply_increment(state, d, grid):
    """Add per-cell increments d[..., 0:3] (ssh at the centre, u on the east
    face, v on the north face, as in t_fields); the outer faces keep their
    boundary values."""
    s = {k: v.copy() for k, v in state.items()}
    s["sshn"] += d[..., 0]
    s["un"][:, 1:-1] += d[:, :-1, 1]
    s["vn"][1:-1, :] += d[:-1, :, 2]
    su, sv = next_interp(s["sshn"], grid["e12t"][0, 0], grid["e12u"][0, 0], grid["e12v"][0, 0])
    s["sshn_u"], s["sshn_v"] = su, sv
    return s


def rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)))


# ---- data ------------------------------------------------------------------------
class Setup:
    """One fine/coarse basin pair."""

    def __init__(self, nc, seed, dx_f=1000.0, rdt_f=20.0, cbfr=0.0015, visc=0.1,
                 radius=(20e3, 200e3), n_mounts=32, height=(10.0, 60.0)):
        self.nc, self.nf = nc, 2 * nc
        self.dx_f, self.rdt_f = dx_f, rdt_f
        self.fine_grid, _ = seamount_grid(self.nf, self.nf, seed, n_mounts=n_mounts, dx=dx_f,
                                          dy=dx_f, radius=radius, height=height)
        self.coarse_grid = coarse_grid(self.fine_grid, 2 * dx_f, 2 * dx_f)
        self.kw = dict(cbfr=cbfr, visc=visc)

    def fine_model(self):
        return Model(self.nf, self.nf, dx=self.dx_f, dy=self.dx_f, rdt=self.rdt_f,
                     grid=self.fine_grid, **self.kw)

    def coarse_model(self):
        return Model(self.nc, self.nc, dx=2 * self.dx_f, dy=2 * self.dx_f, rdt=2 * self.rdt_f,
                     grid=self.coarse_grid, **self.kw)


def generate_pairs(setup, n_snap, spinup_snaps=2):
    """Returns (inputs [S, nc, nc, 4], targets [S, nc, nc, 3]) from one fine
    run of (spinup + n_snap + 1) * K coarse steps."""
    fm, cm = setup.fine_model(), setup.coarse_model()
    snaps = []
    total = spinup_snaps + n_snap + 1
    for s in range(total):
        fm.run(2 * K_STEPS * s + 1, 2 * K_STEPS * (s + 1))
        if s >= spinup_snaps - 1:
            snaps.append((K_STEPS * (s + 1), fm.retrieve()))
    X, Y = [], []
    for (tc, f0), (_, f1) in zip(snaps[:-1], snaps[1:]):
        c0 = restrict_state(f0, setup.fine_grid, setup.coarse_grid)
        cm.set_state(c0)
        cm.run(tc + 1, tc + K_STEPS)
        cp = cm.retrieve()
        truth = t_fields(restrict_state(f1, setup.fine_grid, setup.coarse_grid))
        X.append(features(cp, setup.coarse_grid["ht"]))
        Y.append(truth - t_fields(cp))
    fm.close()
    cm.close()
    return np.array(X), np.array(Y)


# ---- the learned closure ---------------------------------------------------------
class Closure:
    """Normalisation + ConvNet; corrects a coarse state every K steps."""

    def __init__(self, X, Y, seed=0):
        self._init(X.reshape(-1, 4).mean(0), X.reshape(-1, 4).std(0) + 1e-12,
                   Y.reshape(-1, 3).std(0) + 1e-30, seed)

    def _init(self, xm, xs, ys, seed):
        from .closure import ConvNet
        self.net = ConvNet(seed=seed)
        self.xm, self.xs, self.ys = np.asarray(xm), np.asarray(xs), np.asarray(ys)
        # factor on each predicted increment (ssh, u, v) when a model is
        # corrected: 1 applies the network's increment as trained
        self.gain = np.ones(3)

    @classmethod
    def from_stats(cls, xm, xs, ys, seed=0):
        """Input means/scales (4) and target scales (3) given directly."""
        c = cls.__new__(cls)
        c._init(xm, xs, ys, seed)
        return c

    def train_batch(self, xb, yb, n, h, w, wch, lr, want_loss=False):
        """One Adam step on a batch already on the device (xb BF16 features,
        yb FP32 scaled targets, n images of h x w)."""
        self.net.forward(xb, n, h, w)
        loss = self.net.backward(yb, wch, want_loss=want_loss)
        self.net.step(lr=lr)
        return loss

    def _x(self, X):
        n, h, w, _ = X.shape
        xp = np.zeros((n * h * w, 8), dtype=np.float32)
        xp[:, :4] = ((X - self.xm) / self.xs).reshape(-1, 4)
        return self.net.to_bf16(xp)

    def _y(self, Y):
        n, h, w, _ = Y.shape
        yp = np.zeros((n * h * w, 8), dtype=np.float32)
        yp[:, :3] = (Y / self.ys).reshape(-1, 3)
        return self.net.upload(yp)

    def train(self, X, Y, iters=400, batch=8, lr=2e-3, seed=0, log=None):
        rng = np.random.default_rng(seed)
        wch = self.net.upload(np.ones(3, dtype=np.float32))
        n, h, w, _ = X.shape
        losses = []
        for it in range(iters):
            idx = rng.choice(n, size=min(batch, n), replace=False)
            xb, yb = self._x(X[idx]), self._y(Y[idx])
            self.net.forward(xb, len(idx), h, w)
            record = it % 50 == 0 or it == iters - 1
            loss = self.net.backward(yb, wch, want_loss=record)
            self.net.step(lr=lr * (0.5 * (1 + np.cos(np.pi * it / iters))))
            if record:
                losses.append(loss)
                if log:
                    log("closure train it %d loss %.4g" % (it, losses[-1]))
        return losses

    def predict(self, X):
        n, h, w, _ = X.shape
        y = self.net.forward(self._x(X), n, h, w, train=False)
        return y.download()[:, :3].reshape(n, h, w, 3) * self.ys

    def correct_model(self, model):
        """The same correction applied to a Model's device state in place:
        features, network and increments never leave the GPU."""
        from .closure import DevArray
        P = model.jpi * model.jpj
        if getattr(self, "_feat", None) is None or self._feat.shape[0] != P:
            self._feat = DevArray((P, 8), np.int16)
        model.closure_features(self.xm, self.xs, self._feat)
        z = self.net.forward(self._feat, 1, model.jpj, model.jpi, train=False)
        model.closure_apply(z, self.ys * self.gain)

    def correct(self, state, grid):
        d = self.predict(features(state, grid["ht"])[None])[0]
        return apply_increment(state, d, grid), d


def evaluate(setup, closure, n_windows, tol=2.0, judge_every=5):
    """Fine run (the judge) beside a plain coarse run and a coarse+closure run
    from the same restricted start.  Every judge_every windows of K coarse
    steps the closure run is compared with the restricted fine state; when
    its ssh RMS error exceeds tol times the plain run's, the closure is
    switched off (fallback to the plain run's state).  Returns per-window
    ssh RMS errors and whether the fallback fired."""
    fm, cp, cc = setup.fine_model(), setup.coarse_model(), setup.coarse_model()
    fm.run(1, 2 * K_STEPS)
    f = fm.retrieve()
    c0 = restrict_state(f, setup.fine_grid, setup.coarse_grid)
    cp.set_state(c0)
    cc.set_state(c0)
    err_plain, err_clos = [], []
    use = True
    fell_back = False
    for wdx in range(1, n_windows + 1):
        t0 = K_STEPS * wdx
        fm.run(2 * t0 + 1, 2 * (t0 + K_STEPS))
        cp.run(t0 + 1, t0 + K_STEPS)
        cc.run(t0 + 1, t0 + K_STEPS)
        if use:
            closure.correct_model(cc)
        sc = cc.retrieve()
        truth = restrict_state(fm.retrieve(), setup.fine_grid, setup.coarse_grid)
        sp = cp.retrieve()
        err_plain.append(rms(sp["sshn"], truth["sshn"]))
        err_clos.append(rms(sc["sshn"], truth["sshn"]))
        if use and wdx % judge_every == 0 and err_clos[-1] > tol * err_plain[-1]:
            use = False
            fell_back = True
            cc.set_state(sp)
    for m in (fm, cp, cc):
        m.close()
    return np.array(err_plain), np.array(err_clos), fell_back


# ---- the closure at scale: pairs made and used on the device ------------------------
class Trio:
    """One training basin at a time: the fine model, a coarse model holding
    the restricted fine state (the truth) and a coarse model stepping from an
    earlier restricted state (the prediction).  The models are kept and
    given the next basin's grids (load), so device memory is allocated once."""

    def __init__(self, setup):
        self.setup = setup
        self.fine = setup.fine_model()
        self.truth = setup.coarse_model()
        self.pred = setup.coarse_model()
        self.w = 0  # windows of K coarse steps done

    def load(self, setup):
        """The next basin (same sizes): its grids, the fine model at rest."""
        self.setup = setup
        self.fine.set_grid(setup.fine_grid)
        self.fine.set_state(initial_state(setup.fine_grid))
        self.truth.set_grid(setup.coarse_grid)
        self.pred.set_grid(setup.coarse_grid)
        self.w = 0

    def advance(self):
        """Fine model on by one window; truth = R(fine)."""
        self.advance_fine()
        self.advance_truth()

    def advance_fine(self):
        """The fine model's window, queued on its stream (no wait)."""
        w = self.w
        self.fine.run(2 * K_STEPS * w + 1, 2 * K_STEPS * (w + 1))

    def advance_truth(self):
        """truth = R(fine) once the fine window has run."""
        self.truth.restrict_from(self.fine)
        self.w += 1

    def restart_pred(self):
        """Prediction: from R(fine) now, K coarse steps on."""
        self.pred.restrict_from(self.fine)
        self.step_pred()

    def step_pred(self):
        t0 = K_STEPS * self.w
        self.pred.run(t0 + 1, t0 + K_STEPS)

    def close(self):
        for m in (self.fine, self.truth, self.pred):
            m.close()


def pilot_stats(setup, windows=(40, 120, 200)):
    """Normalisation statistics from a few pairs of one basin, made by the
    host path: input mean/scale (4) and target scale (3)."""
    fm, cm = setup.fine_model(), setup.coarse_model()
    X, Y = [], []
    done = 0
    for w in windows:
        fm.run(2 * K_STEPS * done + 1, 2 * K_STEPS * w)
        f0 = fm.retrieve()
        c0 = restrict_state(f0, setup.fine_grid, setup.coarse_grid)
        cm.set_state(c0)
        cm.run(K_STEPS * w + 1, K_STEPS * (w + 1))
        cp_ = cm.retrieve()
        fm.run(2 * K_STEPS * w + 1, 2 * K_STEPS * (w + 1))
        done = w + 1
        truth = t_fields(restrict_state(fm.retrieve(), setup.fine_grid, setup.coarse_grid))
        X.append(features(cp_, setup.coarse_grid["ht"]))
        Y.append(truth - t_fields(cp_))
    fm.close()
    cm.close()
    X, Y = np.array(X), np.array(Y)
    return X.reshape(-1, 4).mean(0), X.reshape(-1, 4).std(0) + 1e-12, Y.reshape(-1, 3).std(0) + 1e-30


def train_online(nc, seeds, windows, group=4, lr=2e-3, spinup=2, bathy=None, iters_per_batch=1,
                 stats=None, log=None, net_seed=0, lr_min_frac=0.05, rollout=1):
    """Closure trained on pairs made on the device: `group` basins advance
    side by side, one window of K coarse steps at a time, and each window's
    `group` pairs (one per basin, written straight into the device batch)
    are one training batch: len(seeds) * windows pairs in all, each used for
    iters_per_batch Adam steps (cosine learning rate).  With rollout > 1 the
    prediction model is restarted from the restricted fine state only every
    `rollout` windows and in between carries on from its own state corrected
    by the network as it is being trained, so that the pairs include states
    the closure itself leads to.  The next group's basins are generated on
    the host while the current group runs.  Returns the closure and a record
    of what was done."""
    import time as _time
    from concurrent.futures import ThreadPoolExecutor
    from .closure import DevArray
    bathy = bathy or {}
    seeds = list(seeds)
    t0 = _time.perf_counter()
    if stats is None:
        stats = pilot_stats(Setup(nc, seeds[0], **bathy))
    clos = Closure.from_stats(*stats, seed=net_seed)
    t_stats = _time.perf_counter() - t0
    P = nc * nc
    xb = DevArray((group * P, 8), np.int16)
    yb = DevArray((group * P, 8), np.float32)
    wch = clos.net.upload(np.ones(3, dtype=np.float32))
    groups = [seeds[k:k + group] for k in range(0, len(seeds), group)]
    total_iters = len(groups) * windows * iters_per_batch
    it = 0
    losses = []
    pairs = 0
    t1 = _time.perf_counter()
    t_wait = 0.0
    trios = []
    pool = ThreadPoolExecutor(max_workers=group)
    # the group's models take their grids side by side (the library's set-up
    # runs on the host without the interpreter lock, one model per thread)
    loader = ThreadPoolExecutor(max_workers=group)

    def make(gs):
        return [pool.submit(Setup, nc, sd, **bathy) for sd in gs]

    pending = make(groups[0])
    for gi in range(len(groups)):
        tw = _time.perf_counter()
        setups = [f.result() for f in pending]
        t_wait += _time.perf_counter() - tw
        if gi + 1 < len(groups):
            pending = make(groups[gi + 1])
        while len(trios) < len(setups):
            trios.append(Trio(setups[len(trios)]))
        active = trios[:len(setups)]
        for f in [loader.submit(tr.load, su) for tr, su in zip(active, setups)]:
            f.result()
        # the group's fine windows are queued on their own streams before
        # any is waited for, so the basins' steps overlap on the device
        for _ in range(spinup):
            for tr in active:
                tr.advance_fine()
            for tr in active:
                tr.advance_truth()
        for tr in active:
            tr.restart_pred()
        for w in range(windows):
            for tr in active:
                tr.advance_fine()
            for tr in active:
                tr.advance_truth()
            for k, tr in enumerate(active):
                tr.pred.closure_features(clos.xm, clos.xs, xb, offset=k * P * 8 * 2)
                tr.truth.closure_target(tr.pred, clos.ys, yb, offset=k * P * 8 * 4)
            pairs += len(active)
            for _ in range(iters_per_batch):
                lr_it = lr * (lr_min_frac + (1 - lr_min_frac) * 0.5 * (1 + np.cos(np.pi * it / total_iters)))
                record = it % 100 == 0 or it == total_iters - 1
                loss = clos.train_batch(xb, yb, len(active), nc, nc, wch, lr_it, want_loss=record)
                if record:
                    losses.append((it, loss))
                    if log:
                        log("closure online it %d/%d pairs %d loss %.4g" % (it, total_iters, pairs, loss))
                it += 1
            for tr in active:
                if (w + 1) % rollout == 0:
                    tr.restart_pred()
                else:
                    clos.correct_model(tr.pred)
                    tr.step_pred()
    pool.shutdown()
    loader.shutdown()
    for tr in trios:
        tr.close()
    t_train = _time.perf_counter() - t1
    xb.close()
    yb.close()
    return clos, {"pairs": pairs, "iterations": it, "batch": group, "losses": losses,
                  "rollout_windows": rollout, "iters_per_batch": iters_per_batch, "lr": lr,
                  "seconds": {"stats": t_stats, "generate_and_train": t_train,
                              "waiting_for_basin_setup": t_wait},
                  "basins": len(seeds), "windows_per_basin": windows, "spinup_windows": spinup}


def evaluate_device(setup, closure, n_windows, tol=2.0, judge_every=5, spinup=2):
    """evaluate() with the restriction on the device: the fine run (the
    judge) beside a plain coarse and a coarse+closure run from the same
    restricted start; every judge_every windows the closure run's ssh RMS
    error against the restricted fine state is compared with the plain
    run's, and above tol times it the closure is switched off (the plain
    run's state taken over).  Returns per-window ssh RMS errors and whether
    the fallback fired."""
    fm, cp_, cc, tm = setup.fine_model(), setup.coarse_model(), setup.coarse_model(), setup.coarse_model()
    fm.run(1, 2 * K_STEPS * spinup)
    cp_.restrict_from(fm)
    cc.restrict_from(fm)
    err_plain, err_clos = [], []
    use, fell_back = True, False
    for wdx in range(spinup, spinup + n_windows):
        t0 = K_STEPS * wdx
        fm.run(2 * t0 + 1, 2 * (t0 + K_STEPS))
        cp_.run(t0 + 1, t0 + K_STEPS)
        cc.run(t0 + 1, t0 + K_STEPS)
        if use:
            closure.correct_model(cc)
        tm.restrict_from(fm)
        truth = tm.retrieve_ssh()
        sp, sc = cp_.retrieve_ssh(), cc.retrieve_ssh()
        err_plain.append(rms(sp, truth))
        err_clos.append(rms(sc, truth))
        if use and (wdx - spinup + 1) % judge_every == 0 and err_clos[-1] > tol * err_plain[-1]:
            use = False
            fell_back = True
            cc.set_state(cp_.retrieve())
    for m in (fm, cp_, cc, tm):
        m.close()
    return np.array(err_plain), np.array(err_clos), fell_back
