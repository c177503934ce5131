"""Benchmark of the dense LU solve path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 32768] [--nb 256]

One step = one solve A x = b of the N x N system already resident in HBM:
PDGETRF + PDGETRS (sla_dgesv), ending at FP64 backward error (the tester's
scaled residual ||b - A x||_inf / (||A||_inf ||x||_inf N eps) < 16 is
checked after the timed steps).  The factorization overwrites A, so every
step restores A from a pristine copy first; the restore is outside the timed
region, which brackets exactly the solve with device synchronisation.

value = (2/3 N^3) / (mean seconds per solve) / 1e12   [TFLOPS-equivalent]

The line also carries
  roofline      the trailing-update GEMM (the dominant kernel): algorithmic
                FP64 flops of all its launches / their summed duration (HIP
                events around each launch on the library's stream, recorded
                during the timed steps) against the MI355X FP64 matrix peak;
  cpu_baseline  the reference ScaLAPACK PDGETRF + PDGETRS (oracle/_ref,
                built from the reference sources, MKL BLAS) on the host cores
                for a bounded N = 16384 sample, same rate definition.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--nb", type=int, default=256)
    ap.add_argument("--cpu-n", type=int, default=16384, help="CPU baseline sample size")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    return ap.parse_args()


def cpu_baseline(n: int) -> dict | None:
    """Reference ScaLAPACK (oracle/_ref/ref_lu) on the host, 1 x 1 grid,
    threaded MKL on the host cores this job may use."""
    ref = ROOT / "oracle" / "_ref" / "ref_lu"
    if not ref.exists():
        return None
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    cores = min(cores, os.cpu_count() or cores)
    env = dict(os.environ, MKL_THREADING_LAYER="GNU", MKL_NUM_THREADS=str(cores),
               OMP_NUM_THREADS=str(cores))
    with tempfile.TemporaryDirectory() as td:
        out = Path(td) / "o.bin"
        # one untimed small run loads MKL's kernels; then the timed sample
        subprocess.run([str(ref), "512", "512", "64", "1", "1", "1", "100", "ones", str(out)],
                       env=env, check=True, capture_output=True)
        proc = subprocess.run([str(ref), str(n), str(n), "256", "1", "1", "1", "100", "ones",
                               str(out)], env=env, check=True, capture_output=True, text=True)
    r = json.loads(proc.stdout.strip().splitlines()[-1])
    t = r["t_getrf"] + r["t_getrs"]
    return {"value": round((2.0 / 3.0) * n ** 3 / t / 1e12, 4), "unit": "TFLOPS",
            "cores": cores, "kind": "reference",
            "sample": f"reference PDGETRF+PDGETRS (oracle/_ref, MKL, {cores} threads, 1x1 grid) "
                      f"on N={n}, NB=256, NRHS=1, PDMATGEN seed 100, b=A*ones; "
                      f"t_getrf={r['t_getrf']:.2f}s t_getrs={r['t_getrs']:.3f}s "
                      f"scaled residual {r['sresid']:.2e}"}


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build "
              f"runs the 1 x 1 BLACS grid on a single MI355X).", file=sys.stderr)
        return 2
    import torch

    from scalapack_amd.device import Handle

    n, nrhs, nb = args.n, args.nrhs, args.nb
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    g = torch.Generator(device=dev)
    g.manual_seed(100)
    # column-major n x n: storage (n cols, n rows); A i.i.d. uniform[-1, 1)
    A0 = torch.rand((n, n), generator=g, device=dev, dtype=torch.float64).mul_(2.0).sub_(1.0)
    B0 = A0.sum(dim=0, keepdim=True).expand(nrhs, n).contiguous()  # b = A * ones
    A = torch.empty_like(A0)
    B = torch.empty_like(B0)
    ipiv = torch.zeros(n, dtype=torch.int32, device=dev)

    h = Handle(nb=nb)
    stream = torch.cuda.current_stream()
    h.set_stream(stream)

    def step():
        return h.dgesv(n, nrhs, A, n, ipiv, B, n)

    for _ in range(args.warmup):
        A.copy_(A0)
        B.copy_(B0)
        torch.cuda.synchronize()
        step()
    torch.cuda.synchronize()

    h.set("profile", 1)
    h.reset_stats()
    times = []
    for _ in range(args.steps):
        A.copy_(A0)
        B.copy_(B0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = step()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        assert info == 0, f"zero pivot at {info}"
    stats = h.stats()
    h.set("profile", 0)

    # backward error of the last solve (FP64 on the GPU)
    X = B
    R = B0.T - A0.T @ X.T
    anorm = A0.abs().sum(dim=0).max().item()       # ||A||_inf: max row sum
    eps = 2.0 ** -53
    sres = (R.abs().amax(dim=0) / (X.T.abs().amax(dim=0) * anorm * eps * n)).max().item()

    ms = 1e3 * sum(times) / len(times)
    value = (2.0 / 3.0) * n ** 3 / (ms / 1e3) / 1e12
    gm = stats["gemm"]
    gemm_tflops = gm["work"] / (gm["ms"] / 1e3) / 1e12 if gm["ms"] > 0 else 0.0
    roof = {"bound": "mfma", "achieved": round(gemm_tflops, 3), "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(gemm_tflops / FP64_PEAK_TFLOPS, 4), "traffic": None,
            "kernel": "dgemm_nn_sub_kernel (trailing update A22 -= L21 U12)",
            "launches_per_step": gm["launches"] // max(1, args.steps),
            "avg_launch_ms": round(gm["ms"] / max(1, gm["launches"]), 4),
            "avg_launch_gflop": round(gm["work"] / max(1, gm["launches"]) / 1e9, 3)}
    breakdown = {k: round(v["ms"] / args.steps, 2) for k, v in stats.items()}

    cpu = None if args.no_cpu else cpu_baseline(args.cpu_n)
    line = {
        "metric": "2N^3/3 TFLOPS-equiv at FP64 backward error: mixed-prec LU+IR vs FP64 LU, N=32k",
        "value": round(value, 3),
        "unit": "TFLOPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: A i.i.d. uniform[-1,1) (torch seed 100), b = A*ones, resident in HBM",
        "config": {"workload": f"dense LU solve PDGESV (1x1 grid), N={n}, NRHS={nrhs}, NB={nb}, "
                               f"plain FP64 path", "n": n, "nrhs": nrhs, "nb": nb,
                   "precision": "fp64"},
        "scaled_residual": sres,
        "kernel_ms_per_step": breakdown,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0 if sres < 16.0 else 1


if __name__ == "__main__":
    sys.exit(main())
