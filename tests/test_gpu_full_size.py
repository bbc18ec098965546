"""GPU against the oracle at the benchmark sizes, on sampled pixels.

The 1024^2 Kerr free-free render and one 512^2 pose of the synchrotron batch are
rendered whole on the GPU (FP64 and mixed precision, wavefront schedule); the oracle
integrates a seeded sample of 192 of their pixels (rtwo_render's subset).  Bars:
FP64 class equal on >= 99 % of the sample and median relative XYZ error <= 1e-5 (the
fixtures' bar); mixed precision class >= 99 % and median <= 1e-3.
"""
import os
import sys

import numpy as np
import pytest

from rtwca_amd import abi, host, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


def _compare(g, o, idx):
    agree = float((g["cls"][idx] == o["cls"][idx]).mean())
    ref = np.abs(o["xyz"][idx]).max(axis=1)
    m = ref > 0
    rel = np.abs(g["xyz"][idx] - o["xyz"][idx]).max(axis=1)[m] / ref[m]
    return agree, (float(np.median(rel)) if m.any() else 0.0), int(m.sum())


@pytest.mark.parametrize("workload", ["kerr_1024", "synch_pose_512"])
def test_sampled_pixels_at_full_size(workload):
    if workload == "kerr_1024":
        w = scenes.config(3, x_pix=1024, schedule=abi.SCHED_WAVEFRONT)
    else:
        w = scenes.config(5, x_pix=512, schedule=abi.SCHED_WAVEFRONT)
        w.cameras = w.cameras[16:17]  # the pose at the top of the inclined orbit
    cam = w.cameras[0]
    npix = cam.x_pix * O.camera_y_pix(cam)
    idx = np.sort(np.random.default_rng(20261203).choice(npix, 192, replace=False))
    o = O.render(w.scene, cam, w.integ, w.lam, nthreads=16, subset=idx)
    ctx = host.Context(0)
    for prec, bar in ((abi.PREC_FP64, 1e-5), (abi.PREC_MIXED, 1e-3)):
        w.integ.precision = prec
        g = ctx.render(w.scene, [cam], w.integ, w.lam)
        assert len(g["cls"]) == npix
        agree, med, nlit = _compare(g, o, idx)
        print("%s %s: class %.4f median rel XYZ %.3g over %d lit sampled pixels" %
              (workload, "fp64" if prec == abi.PREC_FP64 else "mixed", agree, med, nlit), flush=True)
        assert agree >= 0.99
        assert med <= bar
        assert nlit >= 20
