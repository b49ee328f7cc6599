"""The warm-start network's HIP kernels (BF16 MFMA convolutions, FP32
accumulation) against the numpy oracle on the same bf16-rounded inputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(tenet, a):
    from idmx.warmstart import _Buf
    b = _Buf(tenet, a.nbytes)
    b.h2d(a)
    return b


@pytest.mark.parametrize('N,H,W,Cin,Cout', [(2, 8, 8, 8, 16), (1, 16, 12, 16, 1),
                                              (3, 6, 10, 4, 5), (2, 4, 4, 128, 32)])
def test_conv_fwd_bwd_match_oracle(tenet, N, H, W, Cin, Cout):
    from idmx import _lib
    from idmx.warmstart import _Buf, _bf16
    from oracle import nn_ref
    rng = np.random.default_rng(N * 1000 + Cin)
    x = nn_ref.bf16_round(rng.standard_normal((N, H, W, Cin)).astype(np.float32))
    w = (rng.standard_normal((Cout, 9 * Cin)) / np.sqrt(9 * Cin)).astype(np.float32)
    wb = nn_ref.bf16_round(w)
    b = rng.standard_normal(Cout).astype(np.float32)
    h = tenet.handle
    dx_, dw_, db_ = _dev(tenet, _bf16(x)), _dev(tenet, _bf16(w)), _dev(tenet, b)
    dwf = _dev(tenet, w)
    out = _Buf(tenet, N * H * W * Cout * 2)
    out32 = _Buf(tenet, N * H * W * Cout * 4)
    _lib.call('idmx_nn_conv3x3_fwd', h, dx_.ptr, dw_.ptr, db_.ptr, out.ptr, out32.ptr,
              N, H, W, Cin, Cout, 0)
    ref = nn_ref.conv3x3(x, wb, b)
    got = out32.d2h(np.float32, ref.shape)
    assert np.abs(got - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())
    # backward: dL/dout = g (bf16)
    g = nn_ref.bf16_round(rng.standard_normal((N, H, W, Cout)).astype(np.float32))
    dg = _dev(tenet, _bf16(g))
    gw = _Buf(tenet, w.nbytes)
    gb = _Buf(tenet, b.nbytes)
    _lib.call('idmx_nn_conv3x3_bwd_weight', h, dx_.ptr, dg.ptr, gw.ptr, gb.ptr, N, H, W, Cin,
              Cout)
    rw, rb = nn_ref.conv3x3_bwd_weight(x, g)
    assert np.abs(gw.d2h(np.float32, rw.shape) - rw).max() <= 1e-4 * max(1.0, np.abs(rw).max())
    assert np.abs(gb.d2h(np.float32, rb.shape) - rb).max() <= 1e-4 * max(1.0, np.abs(rb).max())
    scratch = _Buf(tenet, w.size * 2)
    dxo = _Buf(tenet, x.size * 2)
    _lib.call('idmx_nn_conv3x3_bwd_data', h, dg.ptr, dwf.ptr, scratch.ptr, dxo.ptr, N, H, W,
              Cin, Cout)
    rd = nn_ref.conv3x3_bwd_data(g, wb, Cin)
    gd = dxo.d2h(np.uint16, x.shape).astype(np.uint32) << 16
    gd = gd.view(np.float32)
    # bf16 output: one rounding of the FP32 accumulation
    assert np.abs(gd - rd).max() <= 8e-3 * max(1.0, np.abs(rd).max())


def test_unet_gradients_match_oracle(tenet):
    """Every layer's weight gradient of the full U-Net (batch 2) against the
    numpy restatement of the same network."""
    from idmx import warmstart as ws
    from oracle import nn_ref
    rng = np.random.default_rng(7)
    net = ws.UNet(tenet, 2, seed=3)
    pts = [(-1.6, 20.0, 128), (-1.5, 40.0, 256)]
    x = np.stack([ws.input_image(G, R, L)[0] for G, R, L in pts]).astype(np.float32)
    tgt = (0.05 * rng.standard_normal((2, ws.S, ws.S))).astype(np.float32)
    net.set_input(ws._bf16(x))
    net.tgt.h2d(tgt)
    net.forward()
    net.loss_and_grad()
    out = net.out32.d2h(np.float32, (2, ws.S, ws.S, 1))
    weights = net.get_weights()
    ref_out, acts = nn_ref.unet_forward(x, weights)
    assert np.abs(out - ref_out).max() <= 2e-2 * max(1e-3, np.abs(ref_out).max())
    net.backward()
    n = 2 * ws.S * ws.S
    dhead = nn_ref.bf16_round((2.0 * (out[..., 0] - tgt) / n)[..., None].astype(np.float32))
    ref = nn_ref.unet_backward(acts, weights, dhead)
    for name, L in net.layers.items():
        gw = L['g'].d2h(np.float32, (L['co'], 9 * L['ci']))
        rw = ref[name][0]
        err = np.linalg.norm(gw - rw) / max(np.linalg.norm(rw), 1e-30)
        assert err < 5e-2, (name, err)
    net.free()


def test_unet_training_reduces_loss(tenet):
    from idmx import warmstart as ws
    rng = np.random.default_rng(0)
    M = 64
    xs = np.stack([ws.input_image(-1.6 + 0.1 * rng.random(), 10 + 30 * rng.random(), 128)[0]
                   for _ in range(M)]).astype(np.float32)
    # a smooth target that depends on the inputs
    tg = (0.1 * np.tanh(5 * xs[..., 2]) * xs[..., 5:6][..., 0]).astype(np.float32)
    net, hist = ws.train(tenet, xs, tg, epochs=30, batch=16, lr=2e-3)
    assert hist[-1] < 0.3 * hist[0]
    net.free()


def test_warm_start_gain_vs_cold_start(tenet):
    """The learned warm start (profile model plus the U-Net's correction,
    populations at rest under the Guo forcing) against the cold start:
    fitted to equilibrated fields of other points, it reaches
    max |dn| < 1e-9 in fewer steps summed over unseen points, at the same
    end states (Laplace pressure within 1e-3, the spread of the droplet's
    two symmetric lattice positions being ~2e-4)."""
    from idmx import sweep
    from idmx import warmstart as ws
    tr = sweep.sample_points(64, 11, Ls=(64, 96), Rfrac=(0.12, 0.28))
    ev = sweep.sample_points(16, 12, Ls=(64, 96), Rfrac=(0.12, 0.28))
    res, f = sweep.run_sweep(tenet, tr, mode='fp64', tol=1e-9)
    conv = [f[i] if res.converged[i] else None for i in range(len(tr))]
    pm = ws.ProfileModel()
    used = pm.fit(tr, conv)
    X, Y, keep = ws.training_set(tr, conv, pm)
    print('training fields: profile %d, net %d' % (used, len(keep)), flush=True)
    assert used >= 32
    net, hist = ws.train(tenet, X, Y, epochs=100, batch=16)
    starts = ws.predict_fields(tenet, net, ev, profile=pm)
    net.free()
    ref, fr = sweep.run_sweep(tenet, ev, mode='fp64', tol=1e-9)
    res_w, fw = sweep.run_sweep(tenet, ev, start=starts, mode='fp64', tol=1e-9)
    ok = [i for i in range(len(ev)) if ref.converged[i] and sweep.is_centred_droplet(fr[i])]
    assert len(ok) >= 9
    s_cold = int(sum(ref.steps[i] for i in ok))
    s_warm = int(sum(res_w.steps[i] for i in ok))
    e_cold = np.median([np.sqrt(np.mean((sweep.cold_start(ev[i]) - fr[i].ravel()) ** 2)) for i in ok])
    e_warm = np.median([np.sqrt(np.mean((starts[i] - fr[i].ravel()) ** 2)) for i in ok])
    print('loss %.3g -> %.3g, start rms error cold %.2e learned %.2e, steps cold %d learned %d '
          '(%.3f)' % (hist[0], hist[-1], e_cold, e_warm, s_cold, s_warm, s_warm / s_cold),
          flush=True)
    assert e_warm < 0.1 * e_cold
    assert s_warm <= 0.95 * s_cold
    for i in ok:
        assert res_w.converged[i]
        dp, dp_ref = sweep.laplace_dp(fw[i], ev[i].G), sweep.laplace_dp(fr[i], ev[i].G)
        assert abs(dp - dp_ref) <= 1e-3 * abs(dp_ref), (i, dp, dp_ref)
