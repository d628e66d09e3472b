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
    ap.add_argument('--pretrain-epochs', type=int, default=3)
    ap.add_argument('--batch', type=int, default=32)
    ap.add_argument('--lr', type=float, default=1e-3)
    ap.add_argument('--seed', type=int, default=11)
    ap.add_argument('--eval', type=int, default=2000)
    ap.add_argument('--eval-seed', type=int, default=7919)
    ap.add_argument('--out', default=WEIGHTS)
    ap.add_argument('--eval-only', action='store_true',
                    help='evaluate the weights in --out on --eval held-out plates and record '
                         'the result in the file')
    a = ap.parse_args(argv)
    log = lambda *s: print(*s, flush=True)
    if a.eval_only:
        net, scales, meta = W.load(a.out)
        evaluate_and_record(net, scales, meta, a, log)
        return
    t0 = time.perf_counter()
    data = W.make_energy_dataset(a.samples, a.seed, log=log)
    log('dataset: %d plates in %.1f s, channel scales %s'
        % (a.samples, time.perf_counter() - t0, data.scales.std))
    net = W.WarmStartNet(n=64, batch=a.batch, seed=a.seed)
    t0 = time.perf_counter()
    hist = W.train_energy(net, data, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed,
                          pretrain_epochs=a.pretrain_epochs, log=log)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log('training: %.1f s' % train_s)
    meta = dict(samples=a.samples, epochs=a.epochs, pretrain_epochs=a.pretrain_epochs,
                batch=a.batch, lr=a.lr, seed=a.seed,
                loss='nodal MSE pre-training, then the energy norm of the Galerkin-scaled '
                     'guess', loss_history=hist, train_s=train_s)
    scales = data.scales
    W.save(net, scales, a.out, meta)
    del data
    torch.cuda.empty_cache()
    if a.eval:
        evaluate_and_record(net, scales, meta, a, log)
    else:
        print(json.dumps(meta), flush=True)


def evaluate_and_record(net, scales, meta, a, log):
    """Zero against warm start on held-out plates; the result, and the
    policy it implies (use the guess only if it saved at least 1 % of the
    block-Jacobi PCG iterations), go into the weights file."""
    res = W.evaluate(net, scales, a.eval, seed=a.eval_seed, log=log)
    meta['eval_seed'] = a.eval_seed
    meta['eval'] = res['summary']
    meta['eval_multigrid'] = res['multigrid']
    meta['use_guess'] = bool(res['summary']['pcg_iteration_ratio'] >= 1.01)
    W.save(net, scales, a.out, meta)
    print(json.dumps(meta), flush=True)


if __name__ == '__main__':
    main()
