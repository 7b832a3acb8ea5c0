"""Write tests/golden/geoclaw_step_vectors.npz from the reference GeoClaw
Fortran step compiled by oracle/Makefile (target ``ref``) from the reference
sources.  The reference's own tests hold no numerical vectors for this path,
so these fixtures are the golden data: inputs and the reference's outputs for
each routine on the path, a single-patch update, a 2x2 brick step, and the
radial dam break run (dt sequence, CFL, mass, and a digest of the final
state).

    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import Brick, CLib, Params, RefLib  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "geoclaw_step_vectors.npz")
DRYTOL, G = 1e-3, 9.81


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def rand_slice(rng, n, mode):
    q = np.zeros((n, 3))
    if mode == "wet":
        h = 0.5 + rng.random(n) * 3
    elif mode == "front":
        h = np.where(np.arange(n) < n // 2, 1 + rng.random(n), 0.0)
    elif mode == "patchy":
        h = rng.random(n) * 2
        h[rng.random(n) < 0.4] = 0.0
        h[rng.random(n) < 0.1] = DRYTOL
    else:  # shallow
        h = rng.random(n) * 5e-3
    q[:, 0] = h
    q[:, 1] = rng.normal(0, 1.5, n) * h
    q[:, 2] = rng.normal(0, 1.5, n) * h
    return q


def main():
    ref = RefLib()
    p = Params()
    ref.set_params(p)
    rng = np.random.default_rng(20240601)
    out = {}

    # riemanntype
    cases = []
    for _ in range(400):
        hL, hR = rng.choice([0.0, 5e-4, 1e-3, 0.3, 2.0, 8.0], 2) * (0.5 + rng.random(2))
        uL, uR = rng.normal(0, 3, 2)
        cases.append((hL, hR, uL, uR))
    cases = np.array(cases)
    rt_out = np.zeros((len(cases), 3))
    rt_rare = np.zeros((len(cases), 2), dtype=np.int32)
    for k, c in enumerate(cases):
        o, r = ref.riemanntype(*c, DRYTOL, G)
        rt_out[k], rt_rare[k] = o, r
    out.update(rt_in=cases, rt_out=rt_out, rt_rare=rt_rare)

    # riemann_aug_JCP on consistent states (as rpn2 prepares them)
    jin, jsw, jfw = [], [], []
    while len(jin) < 300:
        hL, hR = 0.05 + rng.random(2) * 4
        uL, uR = rng.normal(0, 2, 2)
        vL, vR = rng.normal(0, 1, 2)
        bL, bR = rng.normal(0, 0.5, 2)
        phiL = 0.5 * G * hL * hL + (hL * uL) ** 2 / hL
        phiR = 0.5 * G * hR * hR + (hR * uR) ** 2 / hR
        sL, sR = uL - np.sqrt(G * hL), uR + np.sqrt(G * hR)
        uh = (np.sqrt(G * hL) * uL + np.sqrt(G * hR) * uR) / (np.sqrt(G * hR) + np.sqrt(G * hL))
        ch = np.sqrt(G * 0.5 * (hR + hL))
        vin = np.array([hL, hR, hL * uL, hR * uR, hL * vL, hR * vR, bL, bR, uL, uR, vL, vR, phiL, phiR,
                        0.0, 0.0, min(sL, uh - ch), max(sR, uh + ch)])
        sw, fw = ref.riemann_aug_jcp(vin, DRYTOL, G)
        jin.append(vin); jsw.append(sw); jfw.append(fw)
    out.update(jcp_in=np.array(jin), jcp_sw=np.array(jsw), jcp_fw=np.array(jfw))

    # rpn2 / rpt2 / flux2 slices (mx = 16, mbc = 2)
    mx, mbc = 16, 2
    n = mx + 2 * mbc
    modes = ["wet", "front", "patchy", "shallow"]
    sq, sa, sdir = [], [], []
    r_fw, r_s, r_am, r_ap, r_q = [], [], [], [], []
    t_bm, t_bp, t_as = [], [], []
    f_fm, f_fp, f_gm, f_gp, f_cfl = [], [], [], [], []
    for k in range(24):
        q1d = rand_slice(rng, n, modes[k % 4])
        aux = [np.ascontiguousarray((rng.normal(0, 0.6, n) if k % 3 else np.zeros(n))[:, None]) for _ in range(3)]
        ixy = 1 + (k // 4) % 2
        qq = q1d.copy()
        fw, s, am, ap = ref.rpn2(ixy, mx, mbc, qq, aux[1])
        asdq = rng.normal(0, 1, (n, 3))
        bm1, bp1 = ref.rpt2(ixy, 1, mx, mbc, q1d, aux[0], aux[1], aux[2], asdq)
        bm2, bp2 = ref.rpt2(ixy, 2, mx, mbc, q1d, aux[0], aux[1], aux[2], asdq)
        dt = np.full(n, 0.3)
        fm, fp, gm, gp, cfl = ref.flux2(ixy, mx, mbc, q1d.copy(), dt, aux[0], aux[1], aux[2])
        sq.append(q1d); sa.append(np.concatenate(aux, axis=1)); sdir.append(ixy)
        r_fw.append(fw); r_s.append(s); r_am.append(am); r_ap.append(ap); r_q.append(qq)
        t_as.append(asdq); t_bm.append(np.stack([bm1, bm2])); t_bp.append(np.stack([bp1, bp2]))
        f_fm.append(fm); f_fp.append(fp); f_gm.append(gm); f_gp.append(gp); f_cfl.append(cfl)
    out.update(sl_q=np.array(sq), sl_aux=np.array(sa), sl_ixy=np.array(sdir, dtype=np.int32),
               rpn2_fw=np.array(r_fw), rpn2_s=np.array(r_s), rpn2_amdq=np.array(r_am),
               rpn2_apdq=np.array(r_ap), rpn2_q=np.array(r_q), rpt2_asdq=np.array(t_as),
               rpt2_bm=np.array(t_bm), rpt2_bp=np.array(t_bp), flux2_fm=np.array(f_fm),
               flux2_fp=np.array(f_fp), flux2_gm=np.array(f_gm), flux2_gp=np.array(f_gp),
               flux2_cfl=np.array(f_cfl))

    # limiter and philim
    lw, ls, lout = [], [], []
    for k in range(8):
        w = rng.normal(0, 1, (n, 3, 3))
        w[rng.random(n) < 0.2] = 0.0
        s = rng.normal(0, 1, (n, 3))
        w2 = w.copy()
        ref.limiter(mx, mbc, w2, s, (4, 4, 4) if k % 2 == 0 else (1, 2, 3))
        lw.append(w); ls.append(s); lout.append(w2)
    ph_in = np.array([(a, b, m) for m in range(0, 7) for a, b in
                      [(1.0, 2.0), (2.0, -1.0), (0.0, 1.0), (3.0, 0.5), (1.0, 1.0), (0.5, 7.0)]])
    ph_out = np.array([ref.philim(a, b, int(m)) for a, b, m in ph_in])
    out.update(lim_w=np.array(lw), lim_s=np.array(ls), lim_out=np.array(lout), philim_in=ph_in,
               philim_out=ph_out)

    # single patch update (16x16, ghosts given)
    pq, pa, pqo, pc = [], [], [], []
    for k in range(6):
        h = rng.random((20, 20)) * [0.01, 1.0, 4.0][k % 3]
        h[rng.random(h.shape) < [0.0, 0.3, 0.6][k % 3]] = 0.0
        q = np.zeros((20, 20, 3))
        q[..., 0] = h
        q[..., 1] = rng.normal(size=h.shape) * h
        q[..., 2] = rng.normal(size=h.shape) * h
        aux = np.ascontiguousarray(rng.normal(0, [0, 0.4, 1.5][k % 3] + 1e-300, (20, 20, 1)))
        qo = q.copy()
        c = ref.patch_update(16, 16, 2, qo, aux, 0.1, 0.15, 0.0, 0.004)
        pq.append(q); pa.append(aux); pqo.append(qo); pc.append(c)
    out.update(patch_q=np.array(pq), patch_aux=np.array(pa), patch_qout=np.array(pqo),
               patch_cfl=np.array(pc))

    # 2x2 brick of 16x16 patches, one step with the ForestClaw-order ghost fill
    h = 1.0 + rng.random((32, 32)); hu = rng.normal(size=h.shape) * 0.3; hv = rng.normal(size=h.shape) * 0.3
    b = rng.normal(size=h.shape) * 0.2
    br = Brick.from_global(h, hu, hv, b, 2, 2, 2, 0.05, 0.05)
    C = CLib()
    C.set_params(p)
    bc = C.brick_advance(br, 0.0, 0.002, update=ref)
    out.update(brick_in=np.stack([h, hu, hv, b]), brick_out=br.to_global(), brick_cfl=np.array(bc))

    # radial dam break, 8x8 patches of 32x32, 100 steps
    from geoflood_amd.scenarios import radial_dam_break
    h, hu, hv, b, dx = radial_dam_break(256)
    br = Brick.from_global(h, hu, hv, b, 8, 8, 2, dx, dx)
    rec = C.brick_run(br, 100, 1e-4, 0.9, 1.0, update=ref, nthreads=8)
    fin = br.to_global()
    out.update(dam_dt=rec["dt"], dam_cfl=rec["cfl"], dam_mass=rec["mass"],
               dam_retakes=np.array(rec["retakes"]), dam_final_sub=fin[:, ::16, ::16].copy(),
               dam_final_sha256=np.array(digest(fin)), dam_wet_cells=np.array(int((fin[0] > DRYTOL).sum())))

    np.savez_compressed(OUT, **out)
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
