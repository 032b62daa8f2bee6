"""The learned-warm-start U-Net on the GPU against its CPU reference, and the
warm start inside the pressure solve.

- forward pass: BF16 MFMA kernels vs the torch FP32 CPU network with BF16
  rounding at the same points (oracle/unet_ref.py), relative L2 <= 2e-3;
- gradients of 0.5 mean (y - t)^2 with respect to every weight and bias
  vs torch autograd on the CPU, relative L2 <= 5e-2 per layer (the device
  rounds the back-propagated gradients to BF16 as MFMA inputs);
- Adam steps on one sample reduce its loss;
- in the solver: every solve with the learned guess still ends at the same
  FP64 stopping test, and after training the guess lowers the PCG
  iteration count against the plain start (phi = 0).
"""
import ctypes

import numpy as np
import pytest

import bbconfig
import bbhip

pytestmark = pytest.mark.gpu


def lib():
    L = bbhip.load()
    L.bb_unet_create.argtypes = [ctypes.c_ulonglong]
    L.bb_unet_volume.argtypes = [ctypes.c_int] * 3
    L.bb_unet_forward.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.bb_unet_backward.argtypes = [ctypes.c_void_p]
    L.bb_unet_backward.restype = ctypes.c_double
    L.bb_unet_adam.argtypes = [ctypes.c_float]
    L.bb_unet_get.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.bb_unet_layer_info.argtypes = [ctypes.c_int, ctypes.c_void_p]
    return L


def layer_sizes(L):
    info = (ctypes.c_int * 7)()
    out = []
    for l in range(11):
        L.bb_unet_layer_info(l, info)
        cin, cout, taps, _, _, cinp, coutp = list(info)
        out.append((taps * coutp * cinp, coutp))
    return out


def get_params(L, what_w, what_b):
    ws, bs = [], []
    for l, (nw, nb) in enumerate(layer_sizes(L)):
        w = np.zeros(nw, dtype=np.float32)
        b = np.zeros(nb, dtype=np.float32)
        L.bb_unet_get(l, what_w, w.ctypes.data)
        L.bb_unet_get(l, what_b, b.ctypes.data)
        ws.append(w)
        bs.append(b)
    return ws, bs


@pytest.fixture(scope="module")
def net():
    import torch
    L = lib()
    L.cuda_device_init(0)
    L.bb_unet_create(1234)
    yield L, torch
    L.bb_unet_destroy()


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def test_forward_and_gradients(net):
    L, torch = net
    import unet_ref
    dims = (16, 12, 8)
    nv = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(5)
    x = rng.standard_normal((nv, 2)).astype(np.float32)
    t = rng.standard_normal(nv).astype(np.float32)
    assert L.bb_unet_volume(*dims) == 0
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(t).cuda()
    yd = torch.zeros(nv, device="cuda")
    torch.cuda.synchronize()
    L.bb_unet_forward(xd.data_ptr(), yd.data_ptr())
    L.cuda_block()
    loss = L.bb_unet_backward(td.data_ptr())
    L.cuda_block()
    y = yd.cpu().numpy()
    ws, bs = get_params(L, 0, 1)
    gws, gbs = get_params(L, 2, 3)
    rloss, ry, rgw, rgb = unet_ref.grads(ws, bs, x, t, dims)
    yr = ry.transpose(2, 1, 0).ravel()      # (nz, ny, nx) -> channels-last
    ef = rel(y, yr)
    print("forward rel L2", ef, "loss", loss, rloss)
    assert ef <= 2e-3
    assert abs(loss - rloss) <= 2e-3 * abs(rloss)
    for l in range(11):
        ew = rel(gws[l], rgw[l])
        eb = rel(gbs[l][:len(rgb[l])], rgb[l])
        print(f"layer {l}: dW rel {ew:.2e}  db rel {eb:.2e}")
        assert ew <= 5e-2 and eb <= 5e-2, l


def test_adam_reduces_loss(net):
    L, torch = net
    dims = (16, 16, 16)
    nv = 16 ** 3
    rng = np.random.default_rng(7)
    x = rng.standard_normal((nv, 2)).astype(np.float32)
    # a smooth target the network can fit
    g = np.arange(16) / 16.0
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    t = (np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Z)).transpose(2, 1, 0)
    t = t.ravel().astype(np.float32)
    assert L.bb_unet_volume(*dims) == 0
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(t).cuda()
    yd = torch.zeros(nv, device="cuda")
    torch.cuda.synchronize()
    losses = []
    for _ in range(60):
        L.bb_unet_forward(xd.data_ptr(), yd.data_ptr())
        losses.append(L.bb_unet_backward(td.data_ptr()))
        L.bb_unet_adam(ctypes.c_float(3e-3))
    L.cuda_block()
    print("loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.5 * losses[0]


def solve_iterations(case_dir, kw, parts, nsteps, mode, L):
    """PCG iterations per timestep (all Lamb passes) with a given start:
    'zero' (the reference's phi0 = 0), 'last' (span{previous solution}),
    'hist' (span of the last six solutions), 'pass' (those and the previous
    timestep's solutions of the same Lamb pass, the pass before and the
    pass after), 'learned' (that span and the network's guess)."""
    kw = dict(kw)
    ext = dict(kw.get("extensions") or {})
    ext["pp_warmstart"] = "OFF" if mode == "zero" else "LEARNED"
    kw["extensions"] = ext
    bbconfig.write_case(case_dir, **kw)
    bbconfig.write_parts(case_dir, parts)
    hist0 = ctypes.c_int.in_dll(L, "bb_warm_history").value
    pass0 = ctypes.c_int.in_dll(L, "bb_warm_pass_span").value
    ctypes.c_int.in_dll(L, "bb_warm_use_net").value = \
        1 if mode == "learned" else 0
    ctypes.c_int.in_dll(L, "bb_warm_history").value = \
        1 if mode == "last" else 6
    ctypes.c_int.in_dll(L, "bb_warm_pass_span").value = \
        3 if mode in ("pass", "learned") else 0
    tot = ctypes.c_long.in_dll(L, "bb_pcg_iters_total")
    per_step, resid = [], []
    with bbhip.Simulation(case_dir, kw["n"]) as s:
        for _ in range(nsteps):
            t0 = tot.value
            s.step()
            per_step.append(tot.value - t0)
            resid.append(s.pp_stats()["resid"])
        energy = (ctypes.c_double.in_dll(L, "bb_warm_err_last_sum").value,
                  ctypes.c_double.in_dll(L, "bb_warm_err_hist_sum").value,
                  ctypes.c_double.in_dll(L, "bb_warm_err_guess_sum").value,
                  ctypes.c_int.in_dll(L, "bb_warm_net_solves").value)
    ctypes.c_int.in_dll(L, "bb_warm_use_net").value = 1
    ctypes.c_int.in_dll(L, "bb_warm_history").value = hist0
    ctypes.c_int.in_dll(L, "bb_warm_pass_span").value = pass0
    return per_step, resid, energy


def test_warm_start_in_solver(case_dir):
    """Sedimentation box 32^3 with two spheres: 8 training solves, then
    the learned guess on every solve.  Every solve reaches the FP64
    stopping test.  For each solve that used the network the A-norm error
    of its guess is compared with that of the same span without the
    network's direction (both against the converged solution), and the PCG
    iterations of steps 3-4 with the network must be fewer than without
    it (the last solution alone, the fallback start).  Adding the previous
    timestep's solutions of the same Lamb pass, the pass before and the
    pass after to the span of the last six solutions must take fewer
    iterations than that span alone (measured 1919 against 2079)."""
    kw, parts = bbconfig.sediment(n=(32, 32, 32), nparts=2, a=4.0, seed=7)
    L = lib()
    L.bb_unet_destroy()             # the solver builds its own network
    knobs = {"bb_warm_train_solves": 8, "bb_warm_steps": 20,
             "bb_warm_crop": 32, "bb_warm_net_all_passes": 1}
    saved = {k: ctypes.c_int.in_dll(L, k).value for k in knobs}
    for k, v in knobs.items():
        ctypes.c_int.in_dll(L, k).value = v
    try:
        res = {m: solve_iterations(case_dir, kw, parts, 4, m, L)
               for m in ("zero", "last", "hist", "pass", "learned")}
    finally:
        for k, v in saved.items():
            ctypes.c_int.in_dll(L, k).value = v
    for m, (it, rs, _) in res.items():
        print(f"{m:8s} PCG iterations per step {it}  final resid {rs}")
        assert all(r <= 1e-8 for r in rs)
    z, last, hist, pas, lrn = (sum(res[m][0][2:]) for m in
                               ("zero", "last", "hist", "pass", "learned"))
    el, eh, eg, nsolves = res["learned"][2]
    print("steps 3-4: zero", z, "last", last, "hist", hist, "pass", pas,
          "learned", lrn)
    print(f"solves with the network: {nsolves}; A-norm error^2 summed over "
          f"them: last solution {el:.4e}, earlier solutions {eh:.4e}, "
          f"with the network {eg:.4e}")
    assert last < z
    assert nsolves > 0
    assert eg <= eh * (1 + 1e-6)
    assert lrn < last
    assert pas < hist
