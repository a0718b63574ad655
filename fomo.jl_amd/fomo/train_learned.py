"""Train the learned absorbing boundary and evaluate it against the plain HABC.

    python fomo.jl_amd/fomo/train_learned.py --models 512 --out weights.npy

Training runs on the GPU with the engine's own MLP forward/backward kernels (see learned.py).
Evaluation (--eval) applies the reference's absorption measure (test/verify_patch.jl:76-91, its model,
source and receivers) to plain HABC at nbc 16 and 20, the learned boundary at nbc 16 and plain HABC
at nbc 100, against a boundary-free edge-extended reference.  Prints one JSON line per stage.
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--models", type=int, default=512)
    ap.add_argument("--nx", type=int, default=240)
    ap.add_argument("--nz", type=int, default=160)
    ap.add_argument("--nt", type=int, default=500)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--update-every", type=int, default=25)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--init", default="", help="start from these weights (.npy)")
    ap.add_argument("--out", default="")
    ap.add_argument("--eval", action="store_true")
    ap.add_argument("--eval-nx", type=int, default=301)
    ap.add_argument("--eval-nz", type=int, default=201)
    ap.add_argument("--eval-nt", type=int, default=1500)
    ap.add_argument("--eval-L", type=int, default=270)
    ap.add_argument("--weights", default="", help="evaluate these weights instead of training")
    args = ap.parse_args()

    import numpy as np

    from fomo import learned

    def log(m):
        print(f"[learned] {m}", file=sys.stderr, flush=True)

    if args.weights:
        w = np.load(args.weights, allow_pickle=False)
    else:
        init = np.load(args.init, allow_pickle=False) if args.init else None
        t0 = time.time()
        w, hist = learned.train(n_models=args.models, nx=args.nx, nz=args.nz, nt=args.nt, lr=args.lr,
                                update_every=args.update_every, seed=args.seed, weights=init, log=log)
        k = max(1, len(hist) // 8)
        print(json.dumps({"train": {"models": args.models, "seconds": time.time() - t0,
                                    "loss_first": float(np.mean(hist[:k])), "loss_last": float(np.mean(hist[-k:]))}}),
              flush=True)
        if args.out:
            np.save(args.out, np.asarray(w, np.float32))
    if args.eval:
        nt = args.eval_nt
        vp, rho, sx, sz, rx, rz = learned.patch_geometry(args.eval_nx, args.eval_nz)
        res = learned.reflection_residual(vp, rho, sx=sx, sz=sz, rx=rx, rz=rz, nt=nt, L=args.eval_L,
                                          boundaries=[("habc_nbc16", 16, None), ("learned_nbc16", 16, w),
                                                      ("habc_nbc20", 20, None), ("habc_nbc100", 100, None)])
        print(json.dumps({"eval": f"acoustic2d {args.eval_nx}x{args.eval_nz} nt={nt}, fd8, dh 10 dt 1 ms f0 10 Hz, "
                                  "model/source/receivers of test/verify_patch.jl:76-91", **res}), flush=True)


if __name__ == "__main__":
    main()
