"""The reference-signature shim driven by the two-thread phase harness (the stand-in
for sim_thread_loop render.cpp:357-604 and main.cpp's dispatch) against the checker."""
import os
import subprocess

import numpy as np
import pytest

from lenia_amd import workloads as WL
from lenia_amd import build as B

pytestmark = pytest.mark.gpu

KEYS = {"mode": {"srt": 0, "rt": 1, "adv": 2}, "border": {"torus": 0, "wall": 1, "none": 2},
        "gradient": {"central": 0, "sobel": 1}, "growth": {"none": 0, "rt": 1, "transposition": 2}}


def write_config(w, path):
    cfg = w["cfg"]
    lines = [f"mode {KEYS['mode'][cfg['mode']]}", f"border {KEYS['border'][cfg['border']]}",
             f"sobel_border {KEYS['border'][cfg['sobel_border']]}",
             f"gradient {KEYS['gradient'][cfg['gradient']]}", f"growth {KEYS['growth'][cfg['growth']]}",
             f"mu_vel {int(cfg['mu_vel'])}", f"adv_vel {int(cfg['adv_vel'])}", f"blur {int(cfg['blur'])}",
             f"channels {w['C']}"]
    for k in ("dt", "sigma", "dd", "s_norm", "theta_a", "gta", "sta", "v_damp"):
        lines.append(f"{k} {np.float32(cfg[k])!r}".replace("np.float32(", "").replace(")", ""))
    for c in range(4):
        lines.append(f"asymptotic{c} {int(cfg['asymptotic'][c])}")
        lines.append(f"soft_clip{c} {int(cfg['soft_clip'][c])}")
    for k in w["kernels"]:
        vals = [k["c0"], k["c1"], k["radius"], k["beta_size"]] + [repr(float(np.float32(k[n])))
                                                                  for n in ("r_a", "m", "s", "h")]
        vals += [repr(float(np.float32(b))) for b in k["b"]]
        lines.append("kernel " + " ".join(str(v) for v in vals))
    open(path, "w").write("\n".join(lines) + "\n")


def face_pad_expect(field, P):
    """Padded read-back layout [C][(W+8)(H+8)(W+8)] with torus face pads, zero edges."""
    C, D, H_, W = field.shape
    out = np.zeros((C, D + 8, H_ + 8, W + 8), np.float32)
    out[:, 4:-4, 4:-4, 4:-4] = field
    out[:, 4:-4, 4:-4, 3] = field[:, :, :, -1]
    out[:, 4:-4, 4:-4, W + 4] = field[:, :, :, 0]
    out[:, 4:-4, 3, 4:-4] = field[:, :, -1, :]
    out[:, 4:-4, H_ + 4, 4:-4] = field[:, :, 0, :]
    out[:, 3, 4:-4, 4:-4] = field[:, -1, :, :]
    out[:, D + 4, 4:-4, 4:-4] = field[:, 0, :, :]
    return out


@pytest.mark.parametrize("over", [dict(seed=51), dict(seed=52, asymptotic=[0, 1, 1, 0]),
                                  dict(seed=53, mode="srt", asymptotic=[1, 1, 1, 1])])
def test_harness_through_shim_bitwise(tmp_path, over):
    from oracle import oracle as O
    if not os.path.exists(B.HARNESS):
        pytest.fail("ref_harness not built: run __graft_entry__.build()")
    w = WL.plain32(**over)
    C, n, steps = w["C"], 32 ** 3, 4
    write_config(w, tmp_path / "cfg.txt")
    w["A"].astype(np.float32).tofile(tmp_path / "a.f32")
    r = subprocess.run([B.HARNESS, str(tmp_path / "cfg.txt"), str(tmp_path / "a.f32"), str(steps),
                        str(tmp_path / "out.f32")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(tmp_path / "out.f32", dtype=np.float32)
    P = 40 ** 3
    A, Hh, US = (raw[i * C * n:(i + 1) * C * n] for i in range(3))
    RBB, HRB, USRB = (raw[3 * C * n + i * C * P:3 * C * n + (i + 1) * C * P] for i in range(3))
    ow = O.World(w["shape"], C, w["kernels"], w["cfg"], 32)
    ow.set(O.F_A, w["A"])
    ow.step(steps - 1)
    h_prev, us_prev = ow.get(O.F_H), ow.get(O.F_USUM)
    ow.step(1)
    a_ref, h_ref, us_ref = ow.get(O.F_A), ow.get(O.F_H), ow.get(O.F_USUM)
    for got, ref, name in ((A, a_ref, "A"), (Hh, h_ref, "H"), (US, us_ref, "U_SUM")):
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), name
    shape4 = (C, 32, 32, 32)
    assert np.array_equal(RBB.reshape(C, 40, 40, 40), face_pad_expect(a_ref.reshape(shape4), P))
    # H_RB / U_SUM_RB: interior saved at the next fft_step from the previous step's H,
    # face pads from the current H (lenia.cpp:453-457, transposition.cpp:881-995)
    exp_h = face_pad_expect(h_ref.reshape(shape4), P)
    exp_h[:, 4:-4, 4:-4, 4:-4] = h_prev.reshape(shape4)
    assert np.array_equal(HRB.reshape(C, 40, 40, 40), exp_h)
    exp_u = face_pad_expect(us_ref.reshape(shape4), P)
    exp_u[:, 4:-4, 4:-4, 4:-4] = us_prev.reshape(shape4)
    assert np.array_equal(USRB.reshape(C, 40, 40, 40), exp_u)
