"""Build ``libpf3d.so`` (HIP, gfx950) in-tree with hipcc.

``python -m pyfe3d_amd.build`` compiles every ``csrc/*.hip`` into one shared
library next to this file.  The library exports the C ABI declared in
``include/pf3d.h``.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, 'libpf3d.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('PF3D_ARCH', 'gfx950')

FLAGS = [
    '-O3', '-std=c++17', '-fPIC', '--offload-arch=' + ARCH,
    # no packed FP32 instructions anywhere in device code
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-I' + os.path.join(ROOT, 'include'),
    '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
    '-Wno-unused-command-line-argument',
]


def sources():
    return sorted(glob.glob(os.path.join(HERE, 'csrc', '*.hip')))


def build(verbose=False, jobs=8):
    """Compile each source to an object in parallel, then link."""
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + '.o')
        objs.append(obj)
        cmd = [HIPCC] + FLAGS + ['-c', src, '-o', obj]
        if verbose:
            print(' '.join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = False
    for src, p in procs:
        out, _ = p.communicate()
        msg = out.decode(errors='replace')
        if p.returncode != 0:
            failed = True
            sys.stderr.write('hipcc failed on %s\n%s\n' % (src, msg))
        elif verbose and msg.strip():
            sys.stderr.write(msg)
    if failed:
        raise RuntimeError('HIP build failed')
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=' + ARCH, '-shared', '-fPIC', '-o', tmp] + objs
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
