"""Train the warm-start network on plates solved by the GPU solver.

``python -m pyfe3d_amd.train_warmstart --samples 20000 --epochs 12`` draws
seeded 64 x 64 plates (smooth thickness fields, Gaussian pressure patches,
clamped edges), solves them with the batched FP64 PCG, trains the network
with the library's forward/backward/Adam kernels and writes the parameters
to ``pyfe3d_amd/warmstart_weights.npz``; with ``--eval N`` it then compares
zero-start and warm-start solves on N held-out plates.
"""
import argparse
import json
import os
import sys
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
    ap.add_argument('--eval', type=int, default=0)
    ap.add_argument('--data-rtol', type=float, default=1e-6)
    ap.add_argument('--out', default=WEIGHTS)
    a = ap.parse_args(argv)
    log = lambda *s: print(*s, flush=True)
    t0 = time.perf_counter()
    X, T, sc = W.make_dataset(a.samples, a.seed, rtol=a.data_rtol, log=log)
    log('dataset: %d plates in %.1f s, target std %s' % (a.samples, time.perf_counter() - t0,
                                                         sc.std))
    net = W.WarmStartNet(n=64, batch=a.batch, seed=a.seed)
    t0 = time.perf_counter()
    hist = W.train(net, X, T, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed, log=log)
    torch.cuda.synchronize()
    log('training: %.1f s' % (time.perf_counter() - t0))
    W.save(net, sc, a.out)
    res = dict(samples=a.samples, epochs=a.epochs, loss=hist)
    del X, T
    torch.cuda.empty_cache()
    if a.eval:
        res['eval'] = W.evaluate(net, sc, a.eval, seed=a.seed + 1000, log=log)
    print(json.dumps(res), flush=True)


if __name__ == '__main__':
    main()
