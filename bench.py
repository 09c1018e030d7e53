"""Benchmark: FoF particles/s at 1024^3, batched vs plain GPU, bit-exact groups.

Workload: periodic box of 1000 Mpc/h with 1024^3 particles (float32
pos/vel, u64 IDs), half in 2,048,000 Plummer clumps, b = 0.2 (l =
0.1953125 Mpc/h), Nphalomin 20, 3 Mpc/h ghost shell; synthetic, generated on
the GPU.  A step is one full linking call over device-resident input with
device-resident outputs (ring, labels, halo table):

* plain   -- the whole box as one domain plus its periodic shell
* batched -- the box cut into g^3 subvolumes with shells, linked by one batched
             call (``fof_link_gpu_batch``)

value = 1024^3 owned particles / seconds per batched step.  The plain path,
the speedup, per-phase device times, a kernel roofline and a CPU baseline
(the reference's serial algorithm restated in C, one core, on sub-boxes of
the same realization, each also checked for bit-exact groups against the GPU)
are reported on the same JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def mix_ids(ids):
    """64-bit mix of IDs (for order-free membership fingerprints)."""
    import torch

    z = ids.to(torch.int64)
    z = z * 0x2545F4914F6CDD1D
    z = z ^ (z >> 29)
    z = z * 0x5851F42D4C957F2D
    return z ^ (z >> 32)


def group_fingerprints(recs_u8, n, label, halos_u8, nh, vol_off=None, vol_n=None, halo_off=None):
    """(size, sum of mixed member IDs) of every owned halo, sorted: an
    order-free fingerprint of group membership.  Labels and heads are
    volume-local in a batch; vol_off / vol_n / halo_off map them back."""
    import torch

    dev = label.device
    ids = recs_u8[: n * 32].view(torch.int64).view(n, 4)[:, 0]
    lab = label[:n].to(torch.int64) & 0xFFFFFFFF
    h = halos_u8[: nh * 64].view(torch.int32).view(nh, 16)
    head = h[:, 0].to(torch.int64) & 0xFFFFFFFF
    if vol_off is not None:
        off = torch.tensor(np.asarray(vol_off, np.int64), device=dev)
        cnt = torch.tensor(np.asarray(vol_n, np.int64), device=dev)
        lab = lab + torch.repeat_interleave(off, cnt)
        hcnt = torch.tensor(np.diff(np.asarray(halo_off, np.int64)), device=dev)
        head = head + torch.repeat_interleave(off, hcnt)
    size = h[:, 1].to(torch.int64)
    owned = h[:, 2] == 1
    acc = torch.zeros(n, dtype=torch.int64, device=dev)
    acc.index_add_(0, lab, mix_ids(ids))
    fp = torch.stack([size[owned], acc[head[owned]]], 1)
    fp = fp[torch.argsort(fp[:, 1])]
    return fp


# The 13 forward neighbour cells of the reference sweep (src_v2/FOF.c:122-153).
FWD13 = [(1, 0, 0), (-1, 1, 0), (0, 1, 0), (1, 1, 0)] + [
    (dx, dy, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


BIG_CELL = 256  # dense cells from this many particles: the 256-thread kernel


def candidate_pairs(recs_u8, vols, T):
    """Algorithmic work of the linking pass: the pairs the reference tests
    (src_v2/FOF.c:94-161: every pair inside a cell plus every pair with the
    13 forward neighbours, on the reference grid of :45-65), per volume,
    charged to the kernel that tests them.  Dense cells (at least T
    particles, sub-cells enabled) belong to k_link_dense (256-thread
    workgroups from BIG_CELL particles, one wave below); for a constant
    linking length a pair touching a dense cell is tested by that cell's
    kernel (the lower cell's when both are dense), otherwise by the home
    cell's.  Returns ({"big": .., "wave": .., "sparse": ..}, dense_cells)."""
    import torch

    xyz = recs_u8.view(torch.float32).view(-1, 8)[:, 2:5]
    tot = {"big": 0, "wave": 0, "sparse": 0}
    nd = 0
    for v in vols:
        d = v.domain
        a, n = int(v.offset), int(v.n)
        lo = torch.tensor([d.rmin_buff[k] for k in range(3)], dtype=torch.float64, device=xyz.device)
        lc = torch.tensor([d.lcell[k] for k in range(3)], dtype=torch.float64, device=xyz.device)
        dims = [int(np.ceil((d.rmax_buff[k] - d.rmin_buff[k]) / d.lcell[k]) + 1) for k in range(3)]
        s = d.linklength * (1 - 1e-6) / np.sqrt(3.0)
        m = [max(1, int(np.ceil(d.lcell[k] / s))) if s > 0 else 99 for k in range(3)]
        sub_on = all(m[k] <= 12 and dims[k] * m[k] < 2 ** 21 for k in range(3))
        by_dense = not bool(d.varlink)  # constant l: pairs with a dense cell go to its kernel
        idx = torch.empty(n, dtype=torch.int64, device=xyz.device)
        for c0 in range(0, n, 1 << 26):
            c1 = min(n, c0 + (1 << 26))
            t = torch.floor((xyz[a + c0:a + c1].to(torch.float64) - lo) / lc).to(torch.int64)
            for k in range(3):
                t[:, k].clamp_(0, dims[k] - 1)
            idx[c0:c1] = (t[:, 2] * dims[1] + t[:, 1]) * dims[0] + t[:, 0]
        cnt = torch.bincount(idx, minlength=dims[0] * dims[1] * dims[2]).view(dims[2], dims[1], dims[0])
        del idx
        # class of every cell: 0 sparse, 1 wave, 2 big
        cls = torch.zeros_like(cnt, dtype=torch.int8)
        if sub_on:
            cls[cnt >= T] = 1
            cls[(cnt >= T) & (cnt >= BIG_CELL)] = 2
        nd += int((cls > 0).sum())
        acc = torch.zeros(3, dtype=torch.float64, device=cnt.device)

        def add(pairs, klass):
            for k in range(3):
                acc[k] += pairs[klass == k].sum().to(torch.float64)

        add(cnt * (cnt - 1) // 2, cls)

        def sl(dd, nn):
            return (slice(0, nn - 1), slice(1, nn)) if dd == 1 else \
                   (slice(1, nn), slice(0, nn - 1)) if dd == -1 else (slice(0, nn), slice(0, nn))

        for dx, dy, dz in FWD13:
            (zd, zs), (yd, ys), (xd, xs) = sl(dz, dims[2]), sl(dy, dims[1]), sl(dx, dims[0])
            ca, cb = cnt[zd, yd, xd], cnt[zs, ys, xs]
            ka, kb = cls[zd, yd, xd], cls[zs, ys, xs]
            klass = torch.where(ka > 0, ka, kb) if by_dense else ka
            add(ca * cb, klass)
        tot["sparse"] += int(acc[0].item())
        tot["wave"] += int(acc[1].item())
        tot["big"] += int(acc[2].item())
        del cnt, cls
    return tot, nd


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE and WRITE_SIZE passes, scaled by a
    4 GiB copy), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_hbm.txt")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def roofline_link(stats, L, recs_u8, vols, workload=None):
    """Roofline of the dominant link kernel of a timed call: algorithmic FP32
    flops (8 per candidate pair of the reference sweep) over the kernel's
    device time from the library's HIP events on its stream."""
    ph = dict(zip(L.PHASES, stats.phase_ms))
    pairs, nd = candidate_pairs(recs_u8, vols, stats.dense_threshold)
    cands = [("k_link_dense", "link_dense", "big"), ("k_link_dense_wave", "link_dense_wave", "wave"),
             ("k_link_sparse", "link_sparse", "sparse")]
    kern, phase, key = max(cands, key=lambda c: ph[c[1]])
    ms = ph[phase]
    tf = pairs[key] * 8 / (ms * 1e-3) / 1e12
    traffic, src = profiled_traffic(kern, workload) if workload else (None, None)
    return {
        "kernel": kern,
        "bound": "mfma",
        "compute": "FP32 vector ALU (no matrix work on this path); peak = MI355X FP32 157.3 TFLOP/s",
        "achieved": tf,
        "peak": 157.3,
        "unit": "TFLOP/s",
        "frac": tf / 157.3,
        "traffic": traffic,
        "traffic_source": src,
        "kernel_ms": ms,
        "algorithmic_pairs": pairs[key],
        "flop_per_pair": 8,
        "pairs_by_kernel": {"k_link_dense": pairs["big"], "k_link_dense_wave": pairs["wave"],
                            "k_link_sparse": pairs["sparse"]},
        "kernel_ms_by_kernel": {"k_link_dense": ph["link_dense"],
                                "k_link_dense_wave": ph["link_dense_wave"],
                                "k_link_sparse": ph["link_sparse"]},
        "dense_cells_check": {"counted": nd, "engine": int(stats.ndense)},
        "pair_tests_performed": int(stats.pair_tests),
    }


def run_steps(fn, warmup, steps):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def cpu_baseline(recs_u8, bx, budget_s, ctx, L, O):
    """Reference algorithm (oracle.link_ring: src_v2/FOF.c:44-161 + Befriend,
    serial) on interior sub-boxes of the same realization, one core.  Each
    sub-box is also linked on the GPU and its labels compared bit for bit."""
    import torch

    S, shell = 15.625, bx.shell
    rng = np.random.default_rng(2024)
    xyz = recs_u8.view(torch.float32).view(-1, 8)[:, 2:5]
    done_owned, done_t, nbox, exact = 0, 0.0, 0, True
    uf_t = 0.0
    while done_t < budget_s and nbox < 64:
        c = rng.uniform(S / 2 + shell + 1, bx.box - S / 2 - shell - 1, 3)
        lo, hi = c - S / 2 - shell, c + S / 2 + shell
        # coarse float32 cut on the device, exact double cut on the host
        lo_t = torch.tensor(lo - 1.0, device=xyz.device, dtype=torch.float32)
        hi_t = torch.tensor(hi + 1.0, device=xyz.device, dtype=torch.float32)
        m = ((xyz >= lo_t) & (xyz <= hi_t)).all(1)
        idx = torch.nonzero(m).flatten()
        sub = recs_u8.view(-1, 32)[idx].cpu().numpy().reshape(-1).view(L.PART_DTYPE).copy()
        del m
        pp = sub["Pos"].astype(np.float64)
        sub = sub[((pp >= lo) & (pp <= hi)).all(1)].copy()
        rmin, rmax = c - S / 2, c + S / 2
        lc = [bx.box / bx.pside] * 3
        od = O.make_domain(rmin, rmax, lo, hi, lc, bx.linklength, bx.nphalomin)
        t0 = time.perf_counter()
        nxt, _ = O.link_ring(sub["Pos"], od)
        dt = time.perf_counter() - t0
        t1 = time.perf_counter()
        ulab, _ = O.link_uf(sub["Pos"], od)
        uf_t += time.perf_counter() - t1
        p = sub["Pos"].astype(np.float64)
        owned = int((((p >= rmin) & (p < rmax)).all(1)).sum())
        done_owned += owned
        done_t += dt
        nbox += 1
        dom = L.domain(rmin, rmax, lo, hi, lc, bx.linklength, bx.nphalomin)
        glab = np.empty(len(sub), np.uint32)
        rc, _ = ctx.link(L.particles(sub), dom, L.opts(L.OPT_NO_HALOS), None, glab, None, 0)
        L.check(rc)
        rlab = O.labels_from_ring(nxt)
        exact = exact and bool(np.array_equal(glab, rlab)) and bool(np.array_equal(ulab, rlab))
    return {
        "value": done_owned / done_t if done_t > 0 else None,
        "unit": "owned particles/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{nbox} interior sub-boxes of {S} Mpc/h (+{shell} Mpc/h shell) of the same "
                  f"1024^3 realization, {done_owned} owned particles, {done_t:.1f} s; C restatement "
                  f"of the reference's serial cell sweep + ring merge (the reference itself needs "
                  f"GSL headers absent from the image)",
        "gpu_labels_bit_exact_on_sample": exact,
        "uf_restatement_particles_per_s": done_owned / uf_t if uf_t > 0 else None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pside", type=int, default=1024)
    ap.add_argument("--g", type=int, default=4, help="subvolumes per axis for the batched path")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="box", choices=["box", "precision", "volumes", "lightcone"],
                    help="box: the headline (1024^3, plain vs batched); precision: the same box "
                         "with 1%% of particles in borderline pairs; volumes: independent "
                         "periodic volumes with log-normal counts")
    ap.add_argument("--nvol", type=int, default=4096)
    ap.add_argument("--lc-n", type=int, default=400_000_000, help="lightcone particles")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 4 GiB device-to-device copy first (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the FoF domains shard "
              "by rank with no GPU-to-GPU traffic, so value at N GPUs is N x the 1-GPU value",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import oracle as O
    from cm_halofinder_amd import _lib as L
    from cm_halofinder_amd import workloads as W

    dev = torch.device("cuda:0")
    log("device:", L.device_info(0))
    if args.copy_probe:
        a = torch.empty(1 << 32, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        a.fill_(1)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy probe: 4 GiB device-to-device copy done")
        del a, b
    if args.workload == "volumes":
        return run_volumes(args, torch, L, W, O)
    if args.workload == "lightcone":
        return run_lightcone(args, torch, L, W, O)
    if args.workload == "precision":
        bx = W.precision(pside=args.pside)
    else:
        bx = W.plain_1024() if args.pside == 1024 else W.Box(seed=7, pside=args.pside,
                                                             box=1000.0 * args.pside / 1024)
    n = bx.n
    t0 = time.perf_counter()
    recs = torch.empty(n * 32, dtype=torch.uint8, device=dev)
    L.gen_device(bx.gen(), recs)
    log(f"generated {n} particles in {time.perf_counter() - t0:.1f}s")

    def build(g):
        counts, vols = L.ghosts_device(bx.ghost(g), recs, n)
        tot = int(counts.sum())
        out = torch.empty(tot * 32, dtype=torch.uint8, device=dev)
        counts, vols = L.ghosts_device(bx.ghost(g), recs, n, out, tot)
        return out, tot, vols

    t0 = time.perf_counter()
    D1, N1, V1 = build(1)
    DB, NB, VB = build(args.g)
    log(f"ghost shells: plain {N1}, batched {NB} ({args.g}^3 volumes) in {time.perf_counter() - t0:.1f}s")

    ctx = L.Context(0)
    Nmax = max(N1, NB)
    nxt = torch.empty(Nmax, dtype=torch.int32, device=dev)
    lab = torch.empty(Nmax, dtype=torch.int32, device=dev)
    cap = Nmax // 20 + 16
    hal = torch.empty(cap * 64, dtype=torch.uint8, device=dev)
    hoff = np.zeros(len(VB) + 1, np.uint64)
    flags = L.OPT_OUT_DEVICE | L.OPT_TIMING
    st1, stB = L.Stats(), L.Stats()
    P1, PB = L.particles(D1, n=N1), L.particles(DB, n=NB)
    dom1 = V1[0].domain
    nh1 = [0]

    def step_plain():
        rc, nh = ctx.link(P1, dom1, L.opts(flags), nxt, lab, hal, cap, st1)
        L.check(rc, "plain")
        nh1[0] = nh

    def step_batch():
        rc = ctx.link_batch(PB, VB, L.opts(flags), nxt, lab, hal, cap, hoff, stB)
        L.check(rc, "batched")

    t_plain = run_steps(step_plain, args.warmup, args.steps)
    log(f"plain: {t_plain * 1e3:.1f} ms/step  phases {dict(zip(L.PHASES, [round(x, 2) for x in st1.phase_ms]))}")
    fp1 = group_fingerprints(D1, N1, lab, hal, nh1[0])
    t_batch = run_steps(step_batch, args.warmup, args.steps)
    log(f"batched: {t_batch * 1e3:.1f} ms/step  phases {dict(zip(L.PHASES, [round(x, 2) for x in stB.phase_ms]))}")
    nhB = int(hoff[-1])
    fpB = group_fingerprints(DB, NB, lab, hal, nhB, [int(v.offset) for v in VB], [int(v.n) for v in VB],
                             hoff)
    same_groups = bool(fp1.shape == fpB.shape and torch.equal(fp1, fpB))
    log(f"owned halos plain {fp1.shape[0]} batched {fpB.shape[0]} identical membership: {same_groups}")

    roof = roofline_link(stB, L, DB, VB,
                         "box_1024" if args.workload == "box" and args.pside == 1024 and args.g == 4 else None)
    log("roofline:", roof)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(recs, bx, args.cpu_budget, ctx, L, O)
        log("cpu baseline:", cpu)

    value = n / t_batch
    res = {
        "metric": "FoF particles/sec at 1024^3, batched vs plain GPU, bit-exact groups",
        "workload_kind": args.workload,
        "value": value,
        "unit": "owned particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_batch * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"plain FoF, periodic {bx.box:g} Mpc/h box, {bx.pside}^3 particles, "
                        f"{bx.nclumps} clumps, b=0.2, Nphalomin 20, {bx.shell:g} Mpc/h shell; "
                        f"batched = {args.g}^3 subvolumes",
            "particles_with_shells": {"plain": N1, "batched": NB},
        },
        "plain": {"value": n / t_plain, "ms_per_step": t_plain * 1e3,
                  "phase_ms": {k: round(v, 3) for k, v in zip(L.PHASES, st1.phase_ms)}},
        "batched": {"value": value, "ms_per_step": t_batch * 1e3,
                    "phase_ms": {k: round(v, 3) for k, v in zip(L.PHASES, stB.phase_ms)}},
        "speedup_batched_vs_plain": t_plain / t_batch,
        "groups_identical_plain_vs_batched": same_groups,
        "stats": {"pair_tests": stB.pair_tests, "rechecks": stB.rechecks, "ndense": stB.ndense,
                  "nhalos_owned": stB.nhalos_owned, "ngroups": stB.ngroups},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def run_volumes(args, torch, L, W, O):
    """Independent periodic volumes (log-normal counts, median 6e4, clipped to
    [1e3, 2e6], b ~ U(0.15, 0.25)): one batched call for all of them, and one
    call per volume for per-volume latencies; a sample of volumes is checked
    against the oracle."""
    t0 = time.perf_counter()
    P, vols = W.volumes_host(args.nvol, workers=16)
    owned = int(sum(round(float(np.prod([v.domain.rmax[d] - v.domain.rmin[d] for d in range(3)]) /
                                np.prod([v.domain.lcell[d] for d in range(3)]))) for v in vols))
    log(f"built {len(vols)} volumes, {len(P)} particles with shells in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda:0")
    D = torch.from_numpy(P.view(np.uint8).reshape(-1)).to(dev)
    N = len(P)
    ctx = L.Context(0)
    nxt = torch.empty(N, dtype=torch.int32, device=dev)
    lab = torch.empty(N, dtype=torch.int32, device=dev)
    cap = N // 20 + 16 * len(vols)
    hal = torch.empty(cap * 64, dtype=torch.uint8, device=dev)
    hoff = np.zeros(len(vols) + 1, np.uint64)
    st = L.Stats()
    PB = L.particles(D, n=N)
    flags = L.OPT_OUT_DEVICE | L.OPT_TIMING

    def step():
        L.check(ctx.link_batch(PB, vols, L.opts(flags), nxt, lab, hal, cap, hoff, st), "batch")

    t_batch = run_steps(step, args.warmup, args.steps)
    labs = lab.cpu().numpy().view(np.uint32)
    rng = np.random.default_rng(1)
    exact = True
    for k in rng.choice(len(vols), size=min(12, len(vols)), replace=False):
        v = vols[k]
        a, b = int(v.offset), int(v.offset + v.n)
        d = v.domain
        od = O.make_domain(d.rmin, d.rmax, d.rmin_buff, d.rmax_buff, d.lcell, d.linklength, d.nphalomin)
        exact = exact and bool(np.array_equal(labs[a:b], O.link_uf(P["Pos"][a:b], od)[0]))
    # per-volume latency: one synchronous call per volume on device-resident input
    lat = []
    base = D.data_ptr()
    for v in vols:
        pv = L.particles(D, n=int(v.n))
        pv.base = base + int(v.offset) * 32
        t1 = time.perf_counter()
        rc, _ = ctx.link(pv, v.domain, L.opts(L.OPT_OUT_DEVICE), nxt, lab, hal, cap)
        L.check(rc, "volume")
        lat.append(time.perf_counter() - t1)
    lat = np.array(lat) * 1e3
    res = {
        "metric": "FoF particles/sec, batched independent volumes",
        "workload_kind": "volumes",
        "value": owned / t_batch,
        "unit": "owned particles/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{len(vols)} periodic volumes, log-normal counts (median 6e4, "
                               f"[1e3, 2e6]), b ~ U(0.15,0.25), shells", "owned_particles": owned,
                   "particles_with_shells": N},
        "batched_phase_ms": {k: round(x, 3) for k, x in zip(L.PHASES, st.phase_ms)},
        "per_volume_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                  "max": float(lat.max()), "sum": float(lat.sum())},
        "speedup_batched_vs_one_call_per_volume": float(lat.sum() / 1e3 / t_batch),
        "sampled_volumes_bit_exact": exact,
        "roofline": roofline_link(st, L, D, vols),
        "stats": {"pair_tests": st.pair_tests, "rechecks": st.rechecks, "ndense": st.ndense,
                  "nhalos_owned": st.nhalos_owned, "sparse": st.sparse},
    }
    print(json.dumps(res), flush=True)


def run_lightcone(args, torch, L, W, O):
    """Full-sky lightcone to 3000 Mpc/h (flat LCDM, om 0.31), n(r) ~ (1+z)^-1.5,
    l(r) = 0.2 n(r)^(-1/3) as a table on r^2 (home-cell rule), 64 radial
    shells with radial ghosts linked in one batched call (occupied-cell
    addressing); generated on the device.  The two innermost shells are
    checked against the oracle."""
    t0 = time.perf_counter()
    dev = torch.device("cuda:0")
    lc = W.Lightcone(n=args.lc_n)
    P = W.lightcone_particles_device(lc, dev)
    torch.cuda.synchronize()
    log(f"lightcone: {lc.n} particles in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    D, vols, bounds, vt = W.lightcone_shells_device(lc, P)
    del P
    torch.cuda.synchronize()
    N = D.numel() // 32
    log(f"lightcone: {len(vols)} shells, {N} particles with ghosts in {time.perf_counter() - t0:.1f}s")
    ctx = L.Context(0)
    nxt = torch.empty(N, dtype=torch.int32, device=dev)
    lab = torch.empty(N, dtype=torch.int32, device=dev)
    cap = N // 20 + 16 * len(vols)
    hal = torch.empty(cap * 64, dtype=torch.uint8, device=dev)
    hoff = np.zeros(len(vols) + 1, np.uint64)
    st = L.Stats()
    PB = L.particles(D, n=N)
    flags = L.OPT_OUT_DEVICE | L.OPT_TIMING

    def step():
        L.check(ctx.link_batch(PB, vols, L.opts(flags), nxt, lab, hal, cap, hoff, st), "lightcone")

    t_batch = run_steps(step, args.warmup, args.steps)
    log(f"lightcone: {t_batch * 1e3:.1f} ms/step  phases "
        f"{dict(zip(L.PHASES, [round(x, 2) for x in st.phase_ms]))}")
    labs = lab.cpu().numpy().view(np.uint32)
    r2max, tab = lc.varlink_table()
    ovl = O.varlink([0.0, 0.0, 0.0], r2max, tab)
    exact = True
    for k in (0, 1):
        v = vols[k]
        a, b = int(v.offset), int(v.offset + v.n)
        d = v.domain
        od = O.make_domain(d.rmin, d.rmax, d.rmin_buff, d.rmax_buff, d.lcell, d.linklength, d.nphalomin)
        Qk = D[a * 32:b * 32].cpu().numpy().view(L.PART_DTYPE)
        exact = exact and bool(np.array_equal(labs[a:b], O.link_uf(Qk["Pos"], od, ovl)[0]))
    log(f"lightcone: shells 0-1 bit-exact vs oracle: {exact}")
    res = {
        "metric": "FoF particles/sec, variable linking length lightcone, 64 shells batched",
        "workload_kind": "lightcone",
        "value": lc.n / t_batch,
        "unit": "owned particles/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"full-sky lightcone to {lc.rmax:g} Mpc/h, {lc.n} particles, "
                               f"{lc.nshell} shells, l(r) = 0.2 n(r)^-1/3", "particles_with_ghosts": N},
        "phase_ms": {k: round(x, 3) for k, x in zip(L.PHASES, st.phase_ms)},
        "shells_0_1_bit_exact": exact,
        "stats": {"pair_tests": st.pair_tests, "rechecks": st.rechecks, "ndense": st.ndense,
                  "nhalos_owned": st.nhalos_owned, "sparse": st.sparse, "ncells": st.ncells},
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
