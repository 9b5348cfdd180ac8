"""The learned-closure LES experiment, end to end on one GPU:

1. Lundgren-forced DNS on the fine grid (energy held at 0.5), spun up, then
   `nsnap` snapshots `stride` DNS steps apart, each filtered to |k| <= kcut
   on the LES grid: the filtered state ubar and the exact subgrid stress
   tau_hat (kept in device memory).
2. A-priori training of the closure net (MSE on the deviatoric tau,
   channels normalised) on all but the last `test_frac` of the training
   record's snapshots; the loss on those is reported.
3. A-posteriori training, gradients through the LES steps (the solver's
   hand-written RHS adjoint plus the net's backward kernels) reaching the
   net and the scalar gain of the closure force: first `windows`
   trajectory windows of `window` steps from filtered DNS states (squared
   log error of the shell spectrum against the filtered DNS plus a small
   relative-L2 trajectory term), then `attractor_windows` statistical
   windows that start after `spin` gradient-free LES steps, on the LES's
   own attractor, and match the window-averaged spectrum to the filtered
   DNS's mean spectrum. Last, the closure force's gain and spectral tilt
   (g exp(tilt (|k|/kc - 1/2))) are calibrated by two Gauss-Newton steps so
   that the mean spectrum of a `calib_steps` LES matches the training
   snapshots' mean on shells 5..kcut. The LES evolves exactly the
   filter's modes (|k| <= kcut, a sphere inside the solver's 2/3 cube).
4. Evaluation on held-out data: the DNS continues past the training
   record for `holdout_time` more time units, keeping only the filtered
   spectrum of every snapshot (and the filtered state where the record
   starts). An LES from that first held-out state, with and without the
   closure, is compared with it: time-averaged shell spectrum E(k),
   k = 1..kcut, over the held-out window. The closure is accepted when
   max_k |E_les/E_dns - 1| <= tol on this held-out statistic; otherwise the
   result falls back to the plain DNS. Reported beside it: the in-sample
   statistic (an LES from the first training snapshot against the training
   record, the data the gain and tilt were fitted to), the same LES
   continued to `long_factor` times the held-out window against the
   held-out mean, and the DNS's own record-to-record difference (training
   mean against held-out mean: its sampling noise). GPU time of the DNS and
   of the LES over the same simulated time is measured with HIP events.

The LES forcing (HoldResolvedEnergy) is f = A(u) u with A set at every
stage so that the resolved energy relaxes to the filtered DNS's mean
(time scale `relax`), counting the closure's transfer: the DNS holds its
total energy with the same form of forcing, and the filter commutes with
f = A u. Run: python -m hit3d.les_experiment [--quick] [--out FILE]
(with 3dhit_amd on sys.path)."""

from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from .closure import ClosureNet
from .diagnostics import energy_spectrum, quadratic_sums
from .fields import config_spectrum, random_field
from .forcing import LundgrenForcing
from .grid import make_grid, to_physical
from .integrators import rollout
from .les import (HoldResolvedEnergy, LesStepGraph, check_les_state, filter_state, les_rollout,
                  les_step, sphere_mask, subgrid_stress, velocity_gradient)
from .navierstokes import NSParams, make_ns_rhs


def _log(msg):
    print(f"[les {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def _timed(fn):
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    out = fn()
    e1.record()
    e1.synchronize()
    return out, e0.elapsed_time(e1) / 1e3


def _spectrum(u, grid, kmax):
    """Energy of shells 1..kmax."""
    return energy_spectrum(u, grid)[1][:kmax].astype(np.float64)


def _e_omega(u, grid):
    s = quadratic_sums(u, grid)
    return float(s[0]), float(s[2])


def generate(fine_n, les_n, kcut, nu, dt_dns, stride, spinup_time, nsnap, seed, holdout_time=0.0,
             record_training=True):
    """The DNS, its filtered training record and the held-out record. With
    record_training=False the DNS still runs through the training window in
    the same calls (so the held-out record is the same), without filtering
    or keeping its snapshots."""
    fine = make_grid(fine_n, dtype=torch.float32)
    coarse = make_grid(les_n, dtype=torch.float32)
    rhs = make_ns_rhs(fine)
    params = NSParams(nu=nu, forcing=LundgrenForcing(nu=nu))
    u = random_field(seed, fine, config_spectrum)
    nspin = int(round(spinup_time / dt_dns))
    _log(f"DNS {fine_n}^3 spin-up: {nspin} steps (dt {dt_dns})")
    dns_s, done = 0.0, 0
    chunk = max(1, min(nspin, 500))
    while done < nspin:
        n = min(chunk, nspin - done)
        u, s = _timed(lambda: rollout(rhs, u, params, dt_dns, n, done * dt_dns))
        dns_s += s
        done += n
        E, Om = _e_omega(u, fine)
        _log(f"  spin-up step {done}/{nspin}: E = {E:.4f}, Omega = {Om:.2f}, {1e3 * dns_s / done:.2f} ms/step")
    shape = (nsnap if record_training else 0, 3, les_n, les_n, les_n // 2 + 1)
    ubar = torch.empty(shape, dtype=coarse.complex_dtype, device=coarse.device)
    tau = torch.empty((shape[0], 6) + shape[2:], dtype=coarse.complex_dtype, device=coarse.device)
    coef = np.empty(nsnap)
    e_res = np.empty(nsnap)
    t0 = nspin * dt_dns
    dns_steps_s, stress_s = 0.0, 0.0
    for s in range(nsnap):
        u, sec = _timed(lambda: rollout(rhs, u, params, dt_dns, stride, t0 + s * stride * dt_dns))
        dns_steps_s += sec
        if not record_training:
            continue
        (ub, th), sec = _timed(lambda: subgrid_stress(u, fine, coarse, kcut))
        stress_s += sec
        ubar[s], tau[s] = ub, th
        E, Om = _e_omega(u, fine)
        coef[s] = nu * Om / E                     # Lundgren A = 2 nu Omega / (2 E)
        e_res[s] = _e_omega(ub, coarse)[0]        # resolved (filtered) energy
        if (s + 1) % max(1, nsnap // 20) == 0:
            _log(f"  snapshot {s + 1}/{nsnap}: E = {E:.4f}, A = {coef[s]:.4f}, "
                 f"DNS {1e3 * dns_steps_s / ((s + 1) * stride):.2f} ms/step, stress {1e3 * stress_s / (s + 1):.1f} ms")
    # held-out record: the same DNS continued, the filtered spectrum of every
    # snapshot (no stresses), and the filtered state it starts from
    nhold = int(round(holdout_time / (stride * dt_dns)))
    kmax = int(kcut)
    hold_spectra = np.empty((nhold, kmax))
    hold_start, hold_e = None, np.empty(max(nhold, 1))
    t1 = t0 + nsnap * stride * dt_dns
    for s in range(nhold):
        u, sec = _timed(lambda: rollout(rhs, u, params, dt_dns, stride, t1 + s * stride * dt_dns))
        dns_steps_s += sec
        ub = filter_state(u, fine, coarse, kcut)
        if s == 0:
            hold_start = ub.clone()
        hold_spectra[s] = _spectrum(ub, coarse, kmax)
        hold_e[s] = _e_omega(ub, coarse)[0]
        if (s + 1) % max(1, nhold // 10) == 0:
            _log(f"  held-out snapshot {s + 1}/{nhold}, DNS {1e3 * dns_steps_s / ((nsnap + s + 1) * stride):.2f} ms/step")
    del u
    return {"fine": fine, "coarse": coarse, "ubar": ubar, "tau": tau, "coef": coef, "e_res": e_res,
            "dns_ms_per_step": 1e3 * dns_steps_s / ((nsnap + nhold) * stride),
            "stress_ms": 1e3 * stress_s / max(shape[0], 1), "t_start": t0,
            "hold_start": hold_start, "hold_spectra": hold_spectra, "hold_energy": hold_e[:nhold]}


def _deviatoric(tau):
    """tau_ij - delta_ij tr(tau)/3 (components 00, 01, 02, 11, 12, 22): the
    trace enters the LES only as a gradient, which the projection removes."""
    tr = (tau[0] + tau[3] + tau[5]) / 3.0
    out = tau.clone()
    out[0] -= tr
    out[3] -= tr
    out[5] -= tr
    return out


def train_apriori(net, data, train, test, epochs, lr, seed):
    coarse = data["coarse"]
    V = coarse.Nx * coarse.Ny * coarse.Nz
    rng = np.random.default_rng(seed)
    sig2 = (net.out_scale ** 2)[:, None]

    def loss_of(s, train_step):
        g = velocity_gradient(data["ubar"][s], coarse).reshape(9, V)
        tau = _deviatoric(to_physical(data["tau"][s], coarse).reshape(6, V))
        pred, saved = net.forward_saved(g)
        err = pred - tau
        loss = float(torch.mean(err * err / sig2))
        if train_step:
            net.zero_grad()
            net.backward_saved(saved, (2.0 / (6 * V)) * err / sig2, need_grad_g=False)
            net.adam_step(lr)
        return loss

    hist = []
    for ep in range(epochs):
        tr = [loss_of(int(s), True) for s in rng.permutation(train)]
        te = [loss_of(int(s), False) for s in test]
        hist.append({"epoch": ep + 1, "train_mse": float(np.mean(tr)), "test_mse": float(np.mean(te))})
        _log(f"  a-priori epoch {ep + 1}/{epochs}: normalised MSE train {hist[-1]['train_mse']:.4f}, "
             f"held-out {hist[-1]['test_mse']:.4f}")
    return hist


def _shell_energy(u, grid, kmax):
    """Differentiable energy of shells 1..kmax (torch ops on the state)."""
    shell = grid.shell.reshape(-1).long()
    e = (grid.weights * (u.real ** 2 + u.imag ** 2)).sum(0).reshape(-1)
    out = torch.zeros(int(shell.max()) + 1, dtype=e.dtype, device=e.device).index_add_(0, shell, e)
    return out[1:kmax + 1] / (2.0 * float(grid.Nx * grid.Ny * grid.Nz) ** 2)


def train_aposteriori(net, data, params, dt, train, windows, window, lr, seed, kmax, mask,
                      traj_weight=0.1, lr_gain=1e-2, attractor_windows=0, spin=150, target=None):
    """Two kinds of window, each ending in one Adam step of the net's
    parameters (library backward kernels) and of the closure gain:

    - `windows` trajectory windows from filtered DNS states: loss = mean
      over the window's steps of mean_k log(E_les(k)/E_dns(k))^2 plus
      traj_weight times the relative L2^2 distance to the filtered DNS;
    - then `attractor_windows` statistical windows: `spin` LES steps without
      gradient from a filtered DNS state (so that the window starts on the
      LES's own attractor), then `window` steps with gradient; loss =
      mean_k log(<E_les(k)>_window / target(k))^2, target the filtered DNS
      spectrum averaged over the training snapshots. These correct the
      long-time statistics that short trajectory windows cannot see."""
    coarse = data["coarse"]
    rhs = make_ns_rhs(coarse)
    w = coarse.weights
    rng = np.random.default_rng(seed + 1)
    net.reset_optimizer()                 # fresh Adam moments: the a-priori ones are for another loss
    opt_gain = torch.optim.Adam([net.gain], lr=lr_gain)
    hist = []
    total = windows + attractor_windows
    for i in range(total):
        k0 = int(rng.integers(0, len(train) - window))
        u = data["ubar"][int(train[k0])].clone()
        statistical = i >= windows
        if statistical:
            with torch.no_grad():
                for m in range(spin):
                    u = les_step(u, rhs, params, net, dt, 0.0, mask)
        loss, acc = 0.0, 0.0
        for m in range(1, window + 1):
            u = les_step(u, rhs, params, net, dt, (k0 + m - 1) * dt, mask)
            e_les = _shell_energy(u, coarse, kmax)
            if statistical:
                acc = acc + e_les / window
                continue
            ref = data["ubar"][int(train[k0]) + m]
            num = torch.sum(w * (u - ref).abs() ** 2)
            den = torch.sum(w * ref.abs() ** 2)
            lspec = torch.mean(torch.log(e_les / _shell_energy(ref, coarse, kmax)) ** 2)
            loss = loss + (lspec + traj_weight * num / den) / window
        if statistical:
            loss = torch.mean(torch.log(acc / target) ** 2)
        net.zero_grad()
        opt_gain.zero_grad()
        loss.backward()
        net.adam_step(lr)
        opt_gain.step()
        hist.append(float(loss.detach()))
        if (i + 1) % 10 == 0 or i + 1 == total:
            kind = "attractor" if statistical else "trajectory"
            _log(f"  a-posteriori window {i + 1}/{total} ({kind}): loss {np.mean(hist[-10:]):.4e}, "
                 f"gain {float(net.gain.detach()):.3f}")
    return hist


def gauss_newton(resid, x0, deltas, clamp, iters, on_eval=None):
    """Gauss-Newton on a vector residual with a forward-difference Jacobian:
    per iteration one evaluation per coordinate (x + deltas[j] e_j), then the
    least-squares step, clamped to +-clamp. Every evaluated point is kept;
    returns (best x by rms residual, [(x, r), ...] in evaluation order)."""
    x = np.asarray(x0, dtype=np.float64).copy()
    pts = []

    def ev(z):
        r = np.asarray(resid(z), dtype=np.float64)
        pts.append((z.copy(), r))
        if on_eval is not None:
            on_eval(z, r)
        return r

    r = ev(x)
    for _ in range(iters):
        J = np.empty((r.size, x.size))
        for j, d in enumerate(deltas):
            xp = x.copy()
            xp[j] += d
            J[:, j] = (ev(xp) - r) / d
        dx = np.clip(-np.linalg.lstsq(J, r, rcond=None)[0], -np.asarray(clamp), np.asarray(clamp))
        x = x + dx
        r = ev(x)
    best = min(pts, key=lambda p: float(np.sqrt(np.mean(p[1] ** 2))))
    return best[0], pts


def calibrate_closure(net, data, params, dt, kmax, mask, target, steps=2000, iters=2, every=10,
                      kfit=5):
    """Gauss-Newton (`gauss_newton`) on (log gain, tilt) of the closure force
    g exp(tilt (|k|/kc - 1/2)) so that the LES's mean spectrum matches the
    training snapshots' mean on shells kfit..kmax: residual r_k =
    log(<E_les(k)>/target(k)), <.> over the last three quarters of a
    `steps`-step LES from the first snapshot. The gain sets the overall
    drain, the tilt how it is shared between the middle and the top shells.
    Jacobian steps 0.1 in log gain and 0.3 in tilt, updates clamped to 0.2
    and 0.6; the best evaluated point is kept. Uses training data only."""
    coarse = data["coarse"]
    rhs = make_ns_rhs(coarse)
    kfit = max(1, min(kfit, kmax))
    tgt = target.double().cpu().numpy()[kfit - 1:]

    def resid(x):
        with torch.no_grad():
            net.gain.fill_(math.exp(x[0]))
        net.set_tilt(x[1])
        spectra = []
        les_rollout(data["ubar"][0].clone(), rhs, params, net, dt, steps, mask=mask, every=every,
                    observe=lambda n, u: spectra.append(_spectrum(u, coarse, kmax)))
        E = np.mean(spectra[len(spectra) // 4:], axis=0)[kfit - 1:]
        return np.log(E / tgt)

    hist = []

    def record(x, r):
        hist.append({"gain": math.exp(x[0]), "tilt": float(x[1]), "rms_log_error": float(np.sqrt(np.mean(r * r))),
                     "max_log_error": float(np.abs(r).max())})
        _log(f"  closure calibration: gain {math.exp(x[0]):.4f}, tilt {x[1]:+.3f}, "
             f"rms log error {hist[-1]['rms_log_error']:.4f}, max {hist[-1]['max_log_error']:.4f}")

    x0 = [math.log(float(net.gain.detach())), net.tilt]
    best, _ = gauss_newton(resid, x0, (0.1, 0.3), (0.2, 0.6), iters, on_eval=record)
    with torch.no_grad():
        net.gain.fill_(math.exp(best[0]))
    net.set_tilt(best[1])
    return hist


def evaluate(net, grid, u0, ref_spectra, params, dt, steps, kmax, mask, every=10, long_steps=0):
    """LES of max(steps, long_steps) steps from u0 against a filtered DNS
    record whose row n is the spectrum (shells 1..kmax) at LES step n (row 0
    = u0). The statistic pairs LES step n with row n over the first `steps`
    steps, every `every`-th; the long run's mean spectrum is compared with
    the record's mean. Also the cost of one replayed step."""
    rhs = make_ns_rhs(grid)
    check_les_state(u0, grid)
    spectra = []
    nrun = max(steps, long_steps)

    def obs(n, u):
        spectra.append(_spectrum(u, grid, kmax))
    torch.cuda.synchronize()
    _, sec = _timed(lambda: les_rollout(u0.clone(), rhs, params, net, dt, nrun, observe=obs, every=every,
                                        mask=mask))
    npair = min(steps // every, (ref_spectra.shape[0] - 1) // every)
    ref = ref_spectra[every:every * npair + 1:every].mean(0)
    E = np.mean(spectra[:npair], axis=0)
    err = np.abs(E / ref - 1.0)
    out = {"E_les": E.tolist(), "E_dns_filtered": ref.tolist(), "rel_err": err.tolist(),
           "max_rel_err": float(err.max()), "max_rel_err_k_le_20": float(err[:20].max()),
           "worst_k": int(np.argmax(err)) + 1, "steps": steps, "seconds_observed": sec}
    if long_steps > steps:
        E_long = np.mean(spectra, axis=0)
        err_long = np.abs(E_long / ref_spectra.mean(0) - 1.0)
        out.update({"long_steps": nrun, "E_les_long": E_long.tolist(), "rel_err_long": err_long.tolist(),
                    "max_rel_err_long": float(err_long.max())})
    # step cost: replays of the captured step, without observations (each
    # spectrum synchronises the host) and without the one-off capture
    nt = min(100, steps)
    graph = LesStepGraph(u0, rhs, params, net, dt, mask)

    def replays():
        for _ in range(nt):
            graph.step()
    _, sec_pure = _timed(replays)
    out.update({"ms_per_step": 1e3 * sec_pure / nt, "seconds": sec_pure / nt * steps})
    return out


def save_closure(path, net, target, e_target, config):
    """The trained closure (weights, scales, gain, tilt, clipping) and the
    training record's mean spectrum and resolved energy, as one .npz."""
    st = net.state()
    arrays = {"target_spectrum": np.asarray(target, dtype=np.float64), "e_target": np.float64(e_target),
              "in_scale": st["in_scale"].cpu().numpy(), "out_scale": st["out_scale"].cpu().numpy(),
              "gain": st["gain"].cpu().numpy(), "tilt": np.float64(st["tilt"]),
              "clip_backscatter": np.bool_(st["clip_backscatter"]),
              "config": np.bytes_(json.dumps(config))}
    for i, L in enumerate(st["layers"]):
        arrays[f"w{i}"] = L["w"].cpu().numpy()
        arrays[f"b{i}"] = L["b"].cpu().numpy()
    np.savez_compressed(path, **arrays)


def load_closure(path, grid):
    """(ClosureNet on `grid` with the saved state, target spectrum, resolved
    energy target, config dict) from save_closure's file."""
    z = np.load(path)
    config = json.loads(bytes(z["config"]).decode())
    net = ClosureNet(grid, channels=int(z["w0"].shape[0]), seed=0)
    dev = grid.device
    st = {"layers": [{"w": torch.from_numpy(z[f"w{i}"]).to(dev), "b": torch.from_numpy(z[f"b{i}"]).to(dev)}
                     for i in range(3)],
          "in_scale": torch.from_numpy(z["in_scale"]), "out_scale": torch.from_numpy(z["out_scale"]),
          "gain": torch.from_numpy(z["gain"]).to(dev), "clip_backscatter": bool(z["clip_backscatter"]),
          "tilt": float(z["tilt"])}
    net.load(st)
    return net, np.asarray(z["target_spectrum"]), float(z["e_target"]), config


def run(fine_n=512, les_n=64, kcut=21.0, nu=1 / 2000, dt_dns=2e-3, stride=5, spinup_time=10.0,
        nsnap=2000, holdout_time=40.0, test_frac=0.1, epochs=6, lr=1e-3, windows=30, window=50,
        lr_post=3e-4, attractor_windows=100, spin=150, calib_steps=2000, long_factor=3, tol=0.05,
        channels=32, relax=1.0, seed=0, save=None):
    """The experiment end to end; returns the result dict (and writes the
    trained closure to `save` as .npz when given)."""
    wall0 = time.time()
    data = generate(fine_n, les_n, kcut, nu, dt_dns, stride, spinup_time, nsnap, seed, holdout_time)
    coarse = data["coarse"]
    V = coarse.Nx * coarse.Ny * coarse.Nz
    ntrain = nsnap - int(round(test_frac * nsnap))
    train, test = np.arange(ntrain), np.arange(ntrain, nsnap)
    net = ClosureNet(coarse, channels=channels, seed=seed)
    sub = np.linspace(0, ntrain - 1, min(ntrain, 50)).astype(int)
    gs = torch.stack([velocity_gradient(data["ubar"][s], coarse).reshape(9, V).std(dim=1) for s in sub]).mean(0)
    ts = torch.stack([_deviatoric(to_physical(data["tau"][s], coarse).reshape(6, V)).std(dim=1)
                      for s in sub]).mean(0)
    net.set_scales(1.0 / gs, ts)
    _log(f"closure net: {net.num_parameters()} parameters, a-priori training on {ntrain} snapshots "
         f"({len(test)} held out for its loss)")
    apriori = train_apriori(net, data, train, test, epochs, lr, seed)
    dt = stride * dt_dns
    kmax = int(kcut)
    data["spectra"] = np.stack([_spectrum(data["ubar"][s], coarse, kmax) for s in range(nsnap)])
    half = nsnap // 2
    floor = np.abs(data["spectra"][:half].mean(0) / data["spectra"][half:].mean(0) - 1.0)
    e_target = float(np.mean(data["e_res"][:nsnap]))
    params = NSParams(nu=nu, forcing=HoldResolvedEnergy(e_target, relax))
    target_np = data["spectra"].mean(0)
    target = torch.tensor(target_np, dtype=torch.float64, device=coarse.device)
    _log(f"a-posteriori training: {windows} trajectory + {attractor_windows} attractor windows of "
         f"{window} LES steps (dt {dt}, resolved energy held at {e_target:.4f}); training record's own "
         f"half-vs-half spectrum difference: max {floor.max():.3f}")
    mask = sphere_mask(coarse, kcut)
    train_all = np.arange(nsnap)
    apost = train_aposteriori(net, data, params, dt, train_all, windows, window, lr_post, seed, kmax, mask,
                              attractor_windows=attractor_windows, spin=spin, target=target)
    calib = calibrate_closure(net, data, params, dt, kmax, mask, target,
                              steps=min(calib_steps, nsnap - 1)) if calib_steps > 0 else []
    hold = data["hold_spectra"]
    nhold = hold.shape[0]
    # held-out: LES step n pairs with held-out snapshot n (row 0 = the start state)
    ref_hold = np.concatenate([_spectrum(data["hold_start"], coarse, kmax)[None], hold[1:]], axis=0)
    steps = nhold - 1
    _log(f"evaluation on held-out data: {steps} LES steps from the first held-out snapshot "
         f"({long_factor * steps} in all), with and without the closure; in-sample check over the "
         f"training record")
    with_cl = evaluate(net, coarse, data["hold_start"], ref_hold, params, dt, steps, kmax, mask,
                       long_steps=long_factor * steps)
    without = evaluate(None, coarse, data["hold_start"], ref_hold, params, dt, steps, kmax, mask)
    in_sample = evaluate(net, coarse, data["ubar"][0], data["spectra"], params, dt, nsnap - 1, kmax, mask)
    record_diff = np.abs(target_np / hold.mean(0) - 1.0)
    dns_s = data["dns_ms_per_step"] * 1e-3 * steps * stride
    config = {"fine_n": fine_n, "les_n": les_n, "kcut": kcut, "nu": nu, "dt_dns": dt_dns, "dt_les": dt,
              "stride": stride, "spinup_time": spinup_time, "nsnap": nsnap, "holdout_time": holdout_time,
              "holdout_snapshots": nhold, "apriori_train_snapshots": ntrain, "epochs": epochs,
              "windows": windows, "attractor_windows": attractor_windows, "spin": spin, "window": window,
              "calib_steps": calib_steps, "eval_steps": steps, "channels": channels,
              "parameters": net.num_parameters(), "forcing_les": params.forcing.label, "tolerance": tol}
    accepted = with_cl["max_rel_err"] <= tol
    res = {
        "config": config,
        "training_half_vs_half_rel_diff": floor.tolist(), "training_half_vs_half_max": float(floor.max()),
        "training_vs_holdout_rel_diff": record_diff.tolist(),
        "training_vs_holdout_max": float(record_diff.max()),
        "dns_ms_per_step": data["dns_ms_per_step"], "stress_ms_per_snapshot": data["stress_ms"],
        "apriori": apriori, "aposteriori_loss_first_last": [apost[0], float(np.mean(apost[-10:]))],
        "closure_gain": float(net.gain.detach()), "closure_tilt": net.tilt, "closure_calibration": calib,
        "les": with_cl, "les_no_closure": without, "les_in_sample": in_sample,
        "gpu_seconds": {"dns": dns_s, "les": with_cl["seconds"], "speedup": dns_s / with_cl["seconds"]},
        "accepted": accepted, "acceptance_statistic": "held-out",
        "fallback": "none" if accepted else "plain DNS (closure error above tolerance)",
        "wall_s": time.time() - wall0,
    }
    if save:
        save_closure(save, net, target_np, e_target, config)
    _log(f"E(k) max relative error, k = 1..{kmax}, held out: closure {with_cl['max_rel_err']:.3f} "
         f"(k = {with_cl['worst_k']}; k <= 20: {with_cl['max_rel_err_k_le_20']:.3f}; "
         f"{with_cl.get('max_rel_err_long', float('nan')):.3f} over {with_cl.get('long_steps', steps)} steps), "
         f"no closure {without['max_rel_err']:.3f}; in sample {in_sample['max_rel_err']:.3f}; DNS training vs "
         f"held-out record {record_diff.max():.3f}; GPU seconds DNS {dns_s:.1f} vs LES "
         f"{with_cl['seconds']:.2f} ({res['gpu_seconds']['speedup']:.0f}x)")
    return res


def evaluate_saved(path, windows=(2000,), seed=0, tol=0.05):
    """A saved closure against held-out data regenerated from its
    configuration: the DNS (seed `seed`, the experiment's default) is run
    again through spin-up and the training window in the same calls, then
    the held-out record is taken; an LES from the held-out start state is
    compared with it over each window of LES steps in `windows` (the
    benchmark configuration judges 2000 steps) and over the whole held-out
    record. Beside each: the same LES without the closure, and the held-out
    DNS record's own first-half-versus-second-half difference over that
    window (its sampling noise at that averaging length)."""
    cfg = json.loads(bytes(np.load(path)["config"]).decode())
    kcut, kmax, dt = float(cfg["kcut"]), int(cfg["kcut"]), float(cfg["dt_les"])
    data = generate(cfg["fine_n"], cfg["les_n"], kcut, cfg["nu"], cfg["dt_dns"], cfg["stride"],
                    cfg["spinup_time"], cfg["nsnap"], seed, cfg["holdout_time"], record_training=False)
    coarse = data["coarse"]
    net, target, e_target, _ = load_closure(path, coarse)
    params = NSParams(nu=cfg["nu"], forcing=HoldResolvedEnergy(e_target, 1.0))
    mask = sphere_mask(coarse, kcut)
    hold = data["hold_spectra"]
    ref = np.concatenate([_spectrum(data["hold_start"], coarse, kmax)[None], hold[1:]], axis=0)
    nfull = ref.shape[0] - 1
    res = {"closure": str(path), "config": cfg, "seed": seed, "windows": []}
    for w in sorted({min(int(x), nfull) for x in windows} | {nfull}):
        cl = evaluate(net, coarse, data["hold_start"], ref, params, dt, w, kmax, mask)
        bare = evaluate(None, coarse, data["hold_start"], ref, params, dt, w, kmax, mask)
        rows = ref[10:10 * (w // 10) + 1:10]
        h = len(rows) // 2
        noise = np.abs(rows[:h].mean(0) / rows[h:].mean(0) - 1.0)
        err = np.asarray(cl["rel_err"])
        row = {"les_steps": w, "time_units": w * dt, "max_rel_err": cl["max_rel_err"],
               "worst_k": cl["worst_k"], "max_rel_err_k_ge_3": float(err[2:].max()),
               "rel_err": cl["rel_err"], "no_closure_max_rel_err": bare["max_rel_err"],
               "dns_half_vs_half_max": float(noise.max()), "dns_half_vs_half": noise.tolist(),
               "accepted": bool(cl["max_rel_err"] <= tol)}
        res["windows"].append(row)
        _log(f"held-out window of {w} LES steps ({w * dt:g} time units): closure {cl['max_rel_err']:.3f} "
             f"(k = {cl['worst_k']}; k >= 3: {row['max_rel_err_k_ge_3']:.3f}), no closure "
             f"{bare['max_rel_err']:.3f}; the DNS record's own half-vs-half {noise.max():.3f}")
    return res


def drift(path, time_units=120.0, variants=("as saved", "gain x0.85", "gain x1.15", "cut k < 2.5", "cut k < 4.5"),
          seed=5, windows=((10, 40), (40, 80), (80, 120))):
    """Long-run behaviour of a saved closure: an LES from a random-phase
    field with the saved target spectrum (resolved energy held at its
    mean), run for `time_units` with the closure as saved, with its gain
    scaled, or with its force removed below |k| = c ("cut k < c"); per
    window, the time-averaged spectrum's error against the target."""
    out = []
    les_n = int(json.loads(bytes(np.load(path)["config"]).decode())["les_n"])
    for variant in variants:
        grid = make_grid(les_n, dtype=torch.float32)
        net, target, e_target, config = load_closure(path, grid)
        dt = config["dt_les"]
        if variant.startswith("gain x"):
            with torch.no_grad():
                net.gain.mul_(float(variant[6:]))
        elif variant.startswith("cut k < "):
            net.spectral_shape().mul_((torch.sqrt(grid.k2) >= float(variant[8:])).to(grid.dtype))
        elif variant != "as saved":
            raise ValueError(f"unknown variant {variant!r}")
        kmax = len(target)
        spec = lambda k: np.where((k >= 0.5) & (k <= kmax + 0.5),  # noqa: E731
                                  np.interp(k, np.arange(1, kmax + 1), target), 0.0)
        u = filter_state(random_field(seed, grid, spec), grid, grid, float(config["kcut"]))
        params = NSParams(nu=config["nu"], forcing=HoldResolvedEnergy(e_target, 1.0))
        spectra = []
        les_rollout(u, make_ns_rhs(grid), params, net, dt, int(round(time_units / dt)), every=10,
                    mask=sphere_mask(grid, float(config["kcut"])),
                    observe=lambda n, v: spectra.append(_spectrum(v, grid, kmax)))
        S, per = np.asarray(spectra), 10 * dt
        row = {"variant": variant, "dt_les": dt, "windows": []}
        for a, b in windows:
            if b > time_units + 1e-9:
                continue
            err = S[int(round(a / per)):int(round(b / per))].mean(0) / target - 1.0
            row["windows"].append({"from": a, "to": b, "max_abs_err": float(np.abs(err).max()),
                                   "rel_err": err.tolist()})
        _log(f"{variant}: " + ", ".join(f"[{w['from']},{w['to']}] {w['max_abs_err']:.3f}" for w in row["windows"]))
        out.append(row)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--drift", default=None, metavar="CLOSURE_NPZ",
                    help="instead of the experiment: the long-run drift of a saved closure")
    ap.add_argument("--evaluate", default=None, metavar="CLOSURE_NPZ",
                    help="instead of the experiment: a saved closure against held-out data "
                         "regenerated from its configuration")
    ap.add_argument("--eval-steps", dest="eval_steps", type=int, default=2000,
                    help="LES steps of the judged held-out window (with --evaluate)")
    ap.add_argument("--quick", action="store_true", help="small smoke configuration (128^3 -> 32^3)")
    ap.add_argument("--out", default=None, help="write the result JSON here")
    ap.add_argument("--save", default=None, help="write the trained closure (.npz) here")
    ap.add_argument("--holdout-time", dest="holdout_time", type=float, default=None)
    ap.add_argument("--stride", type=int, default=None, help="DNS steps per snapshot and per LES step")
    ap.add_argument("--calib-steps", dest="calib_steps", type=int, default=None)
    ap.add_argument("--nsnap", type=int, default=None)
    ap.add_argument("--epochs", type=int, default=None)
    ap.add_argument("--windows", type=int, default=None)
    ap.add_argument("--attractor-windows", dest="attractor_windows", type=int, default=None)
    a = ap.parse_args(argv)
    if a.drift or a.evaluate:
        res = drift(a.drift) if a.drift else evaluate_saved(a.evaluate, windows=(a.eval_steps,))
        line = json.dumps(res)
        if a.out:
            with open(a.out, "w") as f:
                f.write(line + "\n")
        print(line, flush=True)
        return
    kw = {}
    if a.quick:
        kw = dict(fine_n=128, les_n=32, kcut=10.0, nu=1 / 500, dt_dns=2e-3, stride=4, spinup_time=4.0,
                  nsnap=200, holdout_time=1.6, epochs=2, windows=6, attractor_windows=4, spin=20,
                  window=10, calib_steps=100)
    kw["save"] = a.save
    for k in ("nsnap", "epochs", "windows", "attractor_windows", "holdout_time", "stride", "calib_steps"):
        if getattr(a, k) is not None:
            kw[k] = getattr(a, k)
    res = run(**kw)
    line = json.dumps(res)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    print(line, flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
