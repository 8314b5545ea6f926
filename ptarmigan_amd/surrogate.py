"""The learned sampler of LMA emission: training on the device, saving and
loading trained models.

    python -m ptarmigan_amd.surrogate [--steps N] [--groups G]

trains the linear- and circular-polarisation models on the GPU (the training
data come from the library's own tables and exact sampler, the network is
trained by its own forward/backward kernels) and writes them next to this file
as surrogate_lp.npz and surrogate_cp.npz, with their training reports.
"""
import argparse
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# The domain of the shipped models: the sampler's amplitude a (sqrt(2) a_rms for
# LP, a_rms for CP) and eta. Electrons of 2-20 GeV against a 0.8 um laser have
# eta0 = 0.024 - 0.24; photon recoil lowers eta, so the box starts at 0.01.
A_RANGE = (0.05, 20.0)
ETA_RANGE = (0.01, 0.3)
# Below this amplitude the exact sampler, evaluating J(n, x, y) by quadrature,
# costs no more than the network (measured on plane-wave showers: surrogate with
# FP32 over exact FP64 0.8-1.0x at a0 <= 2, 1.1x at 3, 1.7-3x at 5, 3.5-9x at
# 10-20): the loaded models sample events from this amplitude up, and events
# below it run on the exact kernels in the same generation.
A_USE_MIN = 3.0
# Below this eta the shipped models' spectra are not chi-square consistent with
# the exact sampler's at 1e5 primaries (a0 = 10 showers, independent seeds: CP
# photon energy p = 3e-26 at eta0 = 0.024 (2 GeV), 3e-5 at 0.06, 0.02 at 0.12;
# LP angle p = 1e-5 at 0.024, 1e-3 at 0.06, 0.25 at 0.12): events below it run on
# the exact kernels in the same generation.
ETA_USE_MIN = 0.1


def default_path(polarization):
    return os.path.join(HERE, "surrogate_lp.npz" if int(polarization) == 0 else "surrogate_cp.npz")


def save(model, path, report=None):
    meta = {"polarization": model.polarization, "valid": model.valid, "a_min": model.a_min, "a_max": model.a_max,
            "eta_min": model.eta_min, "eta_max": model.eta_max}
    if report is not None:
        meta["report"] = report_dict(report)
    np.savez(path, params=np.ctypeslib.as_array(model.params).copy(), meta=json.dumps(meta))


def load(path, a_use_min=A_USE_MIN, eta_use_min=ETA_USE_MIN):
    """The model at `path` and its metadata, to be used from amplitude
    `a_use_min` and eta `eta_use_min` up (0: from its box's lower edges)."""
    from . import _abi
    with np.load(path, allow_pickle=False) as z:
        params = z["params"].astype(np.float32)
        meta = json.loads(str(z["meta"]))
    if params.shape != (_abi.MLP_NPARAM,):
        raise ValueError(f"{path}: expected {_abi.MLP_NPARAM} parameters, found {params.shape}")
    m = _abi.SurrogateModel()
    m.polarization = int(meta["polarization"])
    m.valid = int(meta["valid"])
    m.a_min, m.a_max, m.eta_min, m.eta_max = (float(meta[k]) for k in ("a_min", "a_max", "eta_min", "eta_max"))
    m.a_use_min = float(a_use_min)
    m.eta_use_min = float(eta_use_min)
    C.memmove(m.params, params.ctypes.data, params.nbytes)
    return m, meta


def report_dict(rep):
    return {"loss": list(rep.loss), "val_harmonic_rel": rep.val_harmonic_rel, "val_fraction_abs": rep.val_fraction_abs,
            "ks_fraction": rep.ks_fraction, "ks_angle": rep.ks_angle,
            "valid": rep.valid, "groups_used": rep.groups_used, "table_rows": rep.table_rows,
            "sampler_rows": rep.sampler_rows, "gen_ms": rep.gen_ms, "train_ms": rep.train_ms}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--groups", type=int, default=65536)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=HERE)
    ap.add_argument("--pol", choices=["lp", "cp", "both"], default="both")
    args = ap.parse_args(argv)
    from .engine import Engine
    eng = Engine(0)
    for pol, name in ((0, "lp"), (1, "cp")):
        if args.pol not in (name, "both"):
            continue
        print(f"training {name}", flush=True)
        model, rep = eng.surrogate_train(pol, groups=args.groups, steps=args.steps, batch=args.batch, lr=args.lr,
                                         a_range=A_RANGE, eta_range=ETA_RANGE, seed=11 + pol)
        path = os.path.join(args.out, f"surrogate_{name}.npz")
        save(model, path, rep)
        print(json.dumps({"model": name, "path": path, **report_dict(rep)}), flush=True)
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
