"""The LGN adjoint kernels against the float64 oracle: on the 1024-cell
layer of the small plain configuration (128x128 px), 100 ms of white-noise
frames (mean 0.5, sigma 0.3), loss = mean squared error of the rates against
the rates under a drifting grating.

  * forward rates within 1e-5 of the float64 rates (relative to their max);
  * the whole gradient (frames and the 6 per-cell parameters) within 1e-4 of
    the float64 gradient in norm;
  * 256 coordinates (128 pixels, 128 parameters) within 1e-3 relative of
    central finite differences (eps = 1e-3) of the float64 loss; where the
    +-eps runs have different activation patterns (a clipped contrast sample,
    a rectified step, a temporal weight changing sign: kinks of the loss) the
    interval is narrowed to 1e-4, then 1e-5."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def run(pkg, orc):
    import torch
    import oracle as O
    from visual_cortex_amd import network as N
    from visual_cortex_amd.lgn_grad import LgnStage
    spec = N.spec_c1()
    T = 800
    lgn = N.lgn_layer(spec)
    print(f"[lgn adjoint] nLGN={spec.nLGN} {spec.width}x{spec.height} T={T}", flush=True)
    st = LgnStage(spec, lgn, T)
    dev = st.dev
    frames = N.white_noise_frames(spec, st.nF, seed=3)
    grating = N.grating_frames(spec, st.nF)
    th0 = st.theta0.clone()
    with torch.no_grad():
        tgt = st.rates(torch.as_tensor(grating, device=dev), th0)
    F = torch.as_tensor(frames, device=dev).requires_grad_(True)
    TH = th0.clone().requires_grad_(True)
    fr = st.rates(F, TH)
    loss = ((fr - tgt) ** 2).mean()
    loss.backward()
    torch.cuda.synchronize()
    tables = dict(SCx=st.SCx.cpu().numpy(), SCy=st.SCy.cpu().numpy(), SW=st.SW.cpu().numpy(),
                  maxConvol=st.maxConvol.cpu().numpy())
    g64 = O.LgnGrad64(orc, spec, lgn, tables, T, N)
    th64 = th0.cpu().numpy().astype(np.float64)
    tgt_np = tgt.cpu().numpy()
    df64, dth64 = g64.grad(frames, th64, tgt_np)
    return dict(g64=g64, frames=frames, th64=th64, tgt=tgt_np, fr=fr.detach().cpu().numpy(),
                loss=float(loss), dF=F.grad.cpu().numpy(), dTH=TH.grad.cpu().numpy(),
                df64=df64, dth64=dth64)


def test_lgn_rates_match_float64(run):
    fr64 = run["g64"].rates(run["frames"], run["th64"])
    err = np.abs(run["fr"] - fr64).max() / np.abs(fr64).max()
    assert err < 1e-5, err
    l64, _ = run["g64"].loss(run["frames"], run["th64"], run["tgt"])
    assert run["loss"] == pytest.approx(l64, rel=1e-5)


def test_lgn_gradient_matches_float64(run):
    for a, b in ((run["dF"], run["df64"]), (run["dTH"], run["dth64"])):
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rel < 1e-4, rel
    for k in range(6):           # every parameter kind on its own
        a, b = run["dTH"][:, k], run["dth64"][:, k]
        assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b), k


def test_lgn_gradient_finite_differences(run):
    g64, frames, th, tgt = run["g64"], run["frames"], run["th64"], run["tgt"]
    dF, dTH = run["dF"], run["dTH"]
    rng = np.random.default_rng(7)
    nz = np.argwhere(run["df64"] != 0)
    coords = [("f", tuple(c)) for c in nz[rng.choice(len(nz), 128, replace=False)]]
    cells = rng.choice(th.shape[0], 128, replace=False)
    coords += [("t", (int(c), int(k))) for c, k in zip(cells, rng.integers(0, 6, 128))]
    base = frames.astype(np.float64)

    def central(kind, idx, e):
        if kind == "f":
            f, ch, y, x = idx
            sub = g64.cells_touching(x, y, ch)
            fp = base.copy(); fp[idx] += e
            fm = base.copy(); fm[idx] -= e
            lp, pp = g64.loss(fp, th, tgt, sub)
            lm, pm = g64.loss(fm, th, tgt, sub)
        else:
            c, k = idx
            tp = th.copy(); tp[c, k] += e
            tm = th.copy(); tm[c, k] -= e
            lp, pp = g64.loss(frames, tp, tgt, [c])
            lm, pm = g64.loss(frames, tm, tgt, [c])
        return (lp - lm) / (2 * e), pp == pm

    # eps = 1e-3 where the +-eps runs have the same activation pattern; where
    # they straddle a kink (a clipped contrast, a rectified step, a temporal
    # weight changing sign) the interval is narrowed to 1e-4, then 1e-5
    used = {1e-3: 0, 1e-4: 0, 1e-5: 0}
    worst, checked = 0.0, 0
    for kind, idx in coords:
        for e in used:
            fd, same = central(kind, idx, e)
            if same:
                break
        else:
            continue
        used[e] += 1
        if kind == "f":
            gv, scale = dF[idx], np.abs(dF).max()
        else:
            gv, scale = dTH[idx], np.abs(dTH[:, idx[1]]).max()
        assert abs(gv - fd) <= 1e-3 * abs(fd) + 1e-6 * scale, (kind, idx, e, gv, fd)
        worst = max(worst, abs(gv - fd) / max(abs(fd), 1e-30))
        checked += 1
    print(f"[lgn adjoint] finite differences: {checked}/256 checked (eps 1e-3: {used[1e-3]}, "
          f"1e-4: {used[1e-4]}, 1e-5: {used[1e-5]}), worst rel {worst:.2e}")
    assert checked >= 250
