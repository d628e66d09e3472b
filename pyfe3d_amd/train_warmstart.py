"""Train the warm-start network on plates solved by the GPU solver.

``python -m pyfe3d_amd.train_warmstart --samples 20000 --epochs 12 --eval 2000``
draws seeded 64 x 64 plates (smooth thickness fields, Gaussian pressure
patches, clamped edges), solves them with the batched banded Cholesky
(every residual checked at 1e-8), trains the network with the library's
forward/backward/Adam kernels on the energy-norm error of its
Galerkin-scaled guess, evaluates zero against warm start on held-out seeds
(block-Jacobi PCG per plate, multigrid CG on a block-diagonal batch) and
writes the parameters, channel scales and the evaluation summary to
``pyfe3d_amd/warmstart_weights.npz``.
"""
import argparse
import json
import os
import time

import torch

from . import warmstart as W

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, 'warmstart_weights.npz')


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--samples', type=int, default=20000)
    ap.add_argument('--epochs', type=int, default=12)
    ap.add_argument('--batch', type=int, default=32)
    ap.add_argument('--lr', type=float, default=1e-3)
    ap.add_argument('--seed', type=int, default=11)
    ap.add_argument('--eval', type=int, default=2000)
    ap.add_argument('--eval-seed', type=int, default=7919)
    ap.add_argument('--out', default=WEIGHTS)
    a = ap.parse_args(argv)
    log = lambda *s: print(*s, flush=True)
    t0 = time.perf_counter()
    data = W.make_energy_dataset(a.samples, a.seed, log=log)
    log('dataset: %d plates in %.1f s, channel scales %s'
        % (a.samples, time.perf_counter() - t0, data.scales.std))
    net = W.WarmStartNet(n=64, batch=a.batch, seed=a.seed)
    t0 = time.perf_counter()
    hist = W.train_energy(net, data, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed,
                          log=log)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log('training: %.1f s' % train_s)
    meta = dict(samples=a.samples, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed,
                loss='energy norm of the Galerkin-scaled guess', loss_history=hist,
                train_s=train_s)
    scales = data.scales
    W.save(net, scales, a.out, meta)
    del data
    torch.cuda.empty_cache()
    if a.eval:
        res = W.evaluate(net, scales, a.eval, seed=a.eval_seed, log=log)
        meta['eval_seed'] = a.eval_seed
        meta['eval'] = res['summary']
        meta['eval_multigrid'] = res['multigrid']
        W.save(net, scales, a.out, meta)
        log(json.dumps(res['summary']))
    print(json.dumps(meta), flush=True)


if __name__ == '__main__':
    main()
