"""Trains the learned depth prior (unet_prior.py) on a GPU.

Data: seeded synthetic scenes (by default 640x480 with 5 views; --width,
--height, --fx, --views; 4-20 plane patches, 0-50 boxes, 0-40% low-texture
area; seeds from --seed0). Every view of a
scene is a reference with the others as sources. Inputs are the plain
RunPatchMatch maps; labels are the scenes' ground-truth depth and
camera-frame normals (the renderer's ray casts), so the network learns to
correct the plain pass where it fails (low texture, occlusion borders).

    python hpm-mvs_amd/train_unet_prior.py --scenes 200 --steps 3000 --out unet_prior.bin

Prints progress lines; the weights file is read by the pipeline's learned
prior (hpm_pipeline_options.learned_prior_path).
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


def scene_samples(hpm, up, seed, rng, args):
    """(features (H, W, 32) BF16, label (H, W, 4) FP32, gt depth, plain depth, depth range) per view"""
    import numpy as np
    import torch

    w, h = args.width, args.height
    cfg = hpm.scene_config(seed=seed, width=w, height=h, fx=args.fx, fy=args.fx, cx=w / 2.0, cy=h / 2.0,
                           num_views=args.views,
                           num_planes=int(rng.integers(4, 21)), num_boxes=int(rng.integers(0, 51)),
                           low_texture_fraction=float(rng.uniform(0.0, 0.4)))
    cams, imgs, gd, gn, low = hpm.render_scene(cfg)
    n = cfg.num_views
    pool = hpm.Pool(cams, [im.astype(np.float32) for im in imgs])
    problems = [(r, sorted([i for i in range(n) if i != r], key=lambda i: (abs(i - r), i))) for r in range(n)]
    plain = hpm.pipeline_options(planar_prior=0, geom_consistency=0)
    hpm.run_pipeline(pool, problems, plain)
    inputs = [pool.state(v) for v in range(n)]
    pool.close()
    out = []
    for v in range(n):
        cam = cams[v]
        R = np.array(cam.R, np.float32).reshape(3, 3)
        dmin, dmax = 0.6 * cam.depth_min, 1.2 * cam.depth_max
        hyp, cost = inputs[v]
        x = up.features(imgs[v], hyp[..., 3], hyp[..., :3] @ R.T, cost, dmin, dmax)
        lab = np.zeros(hyp.shape, np.float32)
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
    ap.add_argument("--views", type=int, default=5)
    args = ap.parse_args()

    import numpy as np
    import torch

    hpm, up = load_pkg()
    rng = np.random.default_rng(args.seed0)
    t0 = time.time()
    train, val = [], []
    for i in range(args.scenes + args.val_scenes):
        samples = scene_samples(hpm, up, args.seed0 + i, rng, args)
        (val if i >= args.scenes else train).extend(samples)
        if (i + 1) % 10 == 0:
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
        x, lab, _, _, _ = train[int(rng.integers(len(train)))]
        h, w, _ = x.shape
        y0 = int(rng.integers(0, h - cs + 1))
        x0 = int(rng.integers(0, w - cs + 1))
        xc = x[y0:y0 + cs, x0:x0 + cs].contiguous()
        lc = lab[y0:y0 + cs, x0:x0 + cs]
        head = up.forward(xc, ws, bs, conv)
        d, n = up.predict(xc, head)
        ok = torch.isfinite(lc).all(-1)
        lz = torch.nan_to_num(lc)
        per = (d - lz[..., 0]).abs() + 0.1 * (1.0 - (n * lz[..., 1:4]).sum(-1))
        loss = (per * ok).sum() / ok.sum().clamp(min=1)
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
    up.save(args.out, ws, bs)
    log(f"saved {args.out}")
    # held-out views: accuracy (within 2% of ground truth) of the plain
    # maps, of the labels, and of the prediction
    with torch.no_grad():
        acc = np.zeros(3)
        for x, lab, gt, plain_d, (dmin, dmax) in val:
            head = up.forward(x, ws, bs, conv)
            d, _ = up.predict(x, head)
            pred = d * (dmax - dmin) + dmin
            acc += [accuracy(torch.from_numpy(plain_d).cuda(), gt),
                    accuracy(lab[..., 0] * (dmax - dmin) + dmin, gt), accuracy(pred, gt)]
        acc /= max(len(val), 1)
        log(f"held-out accuracy (2%): plain {acc[0]:.4f}, ground truth {acc[1]:.4f}, learned {acc[2]:.4f}")


if __name__ == "__main__":
    main()
