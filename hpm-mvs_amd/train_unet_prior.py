"""Trains the learned depth prior (unet_prior.py) on a GPU.

Data: seeded synthetic scenes (by default 640x480 with 4 views; --width,
--height, --fx, --views; 4-20 plane patches, 0-50 boxes, 0-40% low-texture
area; seeds from --seed0). Every view of a scene is a reference with the
others as sources. Inputs are the plain RunPatchMatch maps. Labels
(--labels):
  plain  the final depth and camera-frame normals of the plain method on
         the same scene (plain pass, planar prior with the KNN supplement,
         K = 8, then 3 geometric-consistency passes), so the network learns
         what the prior and the geometric passes make of the plain pass;
  gt     the scene's ground-truth depth and normals (the renderer's ray
         casts).
The head's channel 4 learns the confidence of the proposal: the logit of
its depth lying within 2% of the label (binary cross-entropy).

    python hpm-mvs_amd/train_unet_prior.py --scenes 2000 --steps 6000 --out unet_prior_plain.bin

Prints progress lines; the weights file is read by the pipeline's learned
prior (hpm_pipeline_options.learned_prior).
"""
import argparse
import importlib.util
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))


def load_pkg():
    name = "hpm_mvs_amd"
    if name not in sys.modules:
        spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, "__init__.py"),
                                                      submodule_search_locations=[HERE])
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
    from hpm_mvs_amd import hpm, unet_prior

    return hpm, unet_prior


def log(msg):
    print(msg, flush=True)


def render(hpm, seed, rng, args):
    w, h = args.width, args.height
    cfg = hpm.scene_config(seed=seed, width=w, height=h, fx=args.fx, fy=args.fx, cx=w / 2.0, cy=h / 2.0,
                           num_views=args.views, arc_length=args.arc,
                           num_planes=int(rng.integers(4, 21)), num_boxes=int(rng.integers(0, 51)),
                           low_texture_fraction=float(rng.uniform(0.0, 0.4)))
    return cfg, hpm.render_scene(cfg)


def scene_samples(hpm, up, seed, rendered, args):
    """per view: (the 6 used input channels (H, W, 6) BF16, label (H, W, 4)
    FP32, gt depth, plain depth, depth range)"""
    import numpy as np
    import torch

    cfg, (cams, imgs, gd, gn, low) = rendered
    n = cfg.num_views
    pool = hpm.Pool(cams, [im.astype(np.float32) for im in imgs])
    problems = [(r, sorted([i for i in range(n) if i != r], key=lambda i: (abs(i - r), i))) for r in range(n)]
    plain = hpm.pipeline_options(planar_prior=0, geom_consistency=0)
    hpm.run_pipeline(pool, problems, plain)
    inputs = [pool.state(v) for v in range(n)]
    labels = None
    if args.labels == "plain":
        full = hpm.pipeline_options()
        full.params.knn_k = 8
        full.params.geom_iterations = 3
        hpm.run_pipeline(pool, problems, full)
        labels = [pool.state(v) for v in range(n)]
    pool.close()
    out = []
    for v in range(n):
        cam = cams[v]
        R = np.array(cam.R, np.float32).reshape(3, 3)
        dmin, dmax = 0.6 * cam.depth_min, 1.2 * cam.depth_max
        hyp, cost = inputs[v]
        x = up.features(imgs[v], hyp[..., 3], hyp[..., :3] @ R.T, cost, dmin, dmax)[..., :6]
        lab = np.zeros(hyp.shape, np.float32)
        if labels is not None:
            lhyp = labels[v][0]
            lab[..., 0] = (lhyp[..., 3] - dmin) / (dmax - dmin)
            lab[..., 1:4] = lhyp[..., :3] @ R.T
        else:
            lab[..., 0] = (gd[v] - dmin) / (dmax - dmin)
            lab[..., 1:4] = gn[v] @ R.T
        bad_x = int((~np.isfinite(x)).sum())
        bad_l = int((~np.isfinite(lab)).any(-1).sum())
        if bad_x or bad_l:
            log(f"scene {seed} view {v}: {bad_x} non-finite inputs, {bad_l} non-finite labels (zeroed / masked)")
        x = np.nan_to_num(x, nan=0.0, posinf=0.0, neginf=0.0)
        lab[~np.isfinite(lab).all(-1)] = np.nan
        out.append((torch.from_numpy(x).to(torch.bfloat16).cuda(), torch.from_numpy(lab).cuda(),
                    torch.from_numpy(gd[v]).cuda(), hyp[..., 3].copy(), (dmin, dmax)))
    return out


def expand(x6):
    """the network input (H, W, 32) from its 6 used channels"""
    import torch

    h, w, _ = x6.shape
    x = torch.zeros(h, w, 32, dtype=x6.dtype, device=x6.device)
    x[..., :6] = x6
    return x


def accuracy(depth, gt):
    valid = gt > 0
    return float(((depth - gt).abs() / gt.clamp(min=1e-6) < 0.02)[valid].float().mean())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=200)
    ap.add_argument("--val-scenes", type=int, default=8)
    ap.add_argument("--seed0", type=int, default=1000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--crop", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=os.path.join(HERE, "unet_prior.bin"))
    ap.add_argument("--width", type=int, default=640)
    ap.add_argument("--height", type=int, default=480)
    ap.add_argument("--fx", type=float, default=500.0)
    ap.add_argument("--views", type=int, default=4)
    ap.add_argument("--arc", type=float, default=0.3, help="camera arc length (m)")
    ap.add_argument("--labels", choices=["plain", "gt"], default="plain")
    ap.add_argument("--conf-weight", type=float, default=0.05)
    args = ap.parse_args()

    import concurrent.futures as cf

    import numpy as np
    import torch

    hpm, up = load_pkg()
    rng = np.random.default_rng(args.seed0)
    t0 = time.time()
    train, val = [], []
    total = args.scenes + args.val_scenes
    # the next scene renders on the host while the device labels this one
    with cf.ThreadPoolExecutor(1) as ex:
        fut = ex.submit(render, hpm, args.seed0, rng, args)
        for i in range(total):
            rendered = fut.result()
            if i + 1 < total:
                fut = ex.submit(render, hpm, args.seed0 + i + 1, rng, args)
            samples = scene_samples(hpm, up, args.seed0 + i, rendered, args)
            (val if i >= args.scenes else train).extend(samples)
            if (i + 1) % 50 == 0:
                log(f"data: {i + 1} scenes, {len(train)} train / {len(val)} val views, {time.time() - t0:.0f} s")
    log(f"{up.num_params()} parameters; training on {len(train)} views")

    ws, bs = up.init_params(0)
    conv = up._conv_function()
    opt = torch.optim.Adam(ws + bs, lr=args.lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, args.steps)
    cs = args.crop
    t1 = time.time()
    run = 0.0
    for step in range(args.steps):
        x, lab, _, _, (dmin, dmax) = train[int(rng.integers(len(train)))]
        h, w, _ = x.shape
        y0 = int(rng.integers(0, h - cs + 1))
        x0 = int(rng.integers(0, w - cs + 1))
        xc = expand(x[y0:y0 + cs, x0:x0 + cs]).contiguous()
        lc = lab[y0:y0 + cs, x0:x0 + cs]
        head = up.forward(xc, ws, bs, conv)
        d, n = up.predict(xc, head)
        ok = torch.isfinite(lc).all(-1)
        lz = torch.nan_to_num(lc)
        per = (d - lz[..., 0]).abs() + 0.1 * (1.0 - (n * lz[..., 1:4]).sum(-1))
        # confidence: is the proposed depth within 2% of the label's
        dd = (d - lz[..., 0]).abs() * (dmax - dmin)
        right = (dd < 0.02 * (lz[..., 0] * (dmax - dmin) + dmin)).float().detach()
        bce = torch.nn.functional.binary_cross_entropy_with_logits(head[..., 4], right, reduction="none")
        loss = ((per + args.conf_weight * bce) * ok).sum() / ok.sum().clamp(min=1)
        if step == 0 or not torch.isfinite(loss):
            log(f"step {step}: loss {loss.item():.5f}, head finite {bool(torch.isfinite(head).all())}, "
                f"input finite {bool(torch.isfinite(xc.float()).all())}")
            if not torch.isfinite(loss):
                raise RuntimeError("non-finite loss")
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        sched.step()
        run = 0.98 * run + 0.02 * loss.item() if step else loss.item()
        if (step + 1) % 100 == 0:
            log(f"step {step + 1}: loss {run:.5f}, {time.time() - t1:.0f} s")
    up.save(args.out, ws, bs, confidence=True)
    log(f"saved {args.out}")
    # held-out views: accuracy (within 2% of ground truth) of the plain
    # maps, of the labels, of the prediction, and of the prediction where
    # the confidence is above one half (with that fraction)
    with torch.no_grad():
        acc = np.zeros(5)
        for x6, lab, gt, plain_d, (dmin, dmax) in val:
            x = expand(x6)
            head = up.forward(x, ws, bs, conv)
            d, _ = up.predict(x, head)
            pred = d * (dmax - dmin) + dmin
            conf = head[..., 4] > 0
            sure = conf & (gt > 0)
            acc_sure = float(((pred - gt).abs() / gt.clamp(min=1e-6) < 0.02)[sure].float().mean()) if sure.any() else 0
            acc += [accuracy(torch.from_numpy(plain_d).cuda(), gt),
                    accuracy(lab[..., 0] * (dmax - dmin) + dmin, gt), accuracy(pred, gt), acc_sure,
                    float(conf.float().mean())]
        acc /= max(len(val), 1)
        log(f"held-out accuracy (2%): plain pass {acc[0]:.4f}, labels ({args.labels}) {acc[1]:.4f}, "
            f"learned {acc[2]:.4f}, learned where confident {acc[3]:.4f} (confident fraction {acc[4]:.3f})")


if __name__ == "__main__":
    main()
