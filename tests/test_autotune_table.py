"""Host logic (no GPU): the autotuner's table of stored choices and the
sweep's grouping of systems for the split banded factorisation."""
import json

import numpy as np

from pyfe3d_amd import autotune as AT


def test_keys_and_buckets():
    assert AT.size_bucket(4096) == 4096
    assert AT.size_bucket(8165) == 8192
    assert AT.size_bucket(5000) == 4096
    assert AT.size_bucket(0) == 1
    assert AT.make_key('band_direct', 'quad4', (448, 4096, 8192), 'gfx950') == \
        'band_direct|quad4|448x4096x8192|gfx950'


def test_record_then_lookup_prefers_user_table(tmp_path, monkeypatch):
    shipped = tmp_path / 'shipped.json'
    shipped.write_text(json.dumps({'k|a': {'config': {'ft': 1}}, 'k|b': {'config': {'ft': 2}}}))
    monkeypatch.setenv('PF3D_AUTOTUNE_CACHE', str(tmp_path / 'sub' / 'user.json'))
    assert AT.lookup('k|a', shipped=str(shipped))['config'] == {'ft': 1}
    assert AT.lookup('k|c', shipped=str(shipped)) is None
    assert AT.record('k|a', {'ft': 7}, [({'ft': 7}, 1.25), ({'ft': 1}, 2.5)])
    e = AT.lookup('k|a', shipped=str(shipped))
    assert e['config'] == {'ft': 7} and e['timings'][1] == {'ft': 1, 'ms': 2.5}
    assert AT.lookup('k|b', shipped=str(shipped))['config'] == {'ft': 2}


def test_cache_off_and_unwritable(tmp_path, monkeypatch):
    monkeypatch.setenv('PF3D_AUTOTUNE_CACHE', 'off')
    assert AT.user_path() is None
    assert not AT.record('k', {'ft': 1})
    blocker = tmp_path / 'file'
    blocker.write_text('x')
    monkeypatch.setenv('PF3D_AUTOTUNE_CACHE', str(blocker / 'user.json'))
    assert not AT.record('k', {'ft': 1})  # parent is a file: memo lost, no error


def test_shipped_table_entry_for_the_design_sweep():
    data = json.load(open(AT.SHIPPED))
    e = data['band_direct|quad4|448x4096x8192|gfx950']
    c = e['config']
    assert c['ft'] in (256, 512, 1024) and c['mb'] in (256, 448, 576)
    assert c['st'] in (256, 512, 1024)
    best = min(e['timings'], key=lambda t: t['ms'])
    assert c.get('sft', 0) in (0, 256, 512)
    assert {k: best.get(k, 0) for k in c} == c


def test_shipped_table_entries_for_the_batched_pcg():
    data = json.load(open(AT.SHIPPED))
    for mode in ('fp64', 'mixed'):
        e = data['batch_pcg_%s|quad4|256x8192|gfx950' % mode]
        c = e['config']
        assert c['threads'] in (256, 512, 1024) and c['groups'] > 0
        best = min(e['timings'], key=lambda t: t['ms'])
        assert {k: best[k] for k in c} == c


def test_split_order_of_the_banded_factorisation():
    """Wide bands (> 256) first, then the narrow ones, largest systems
    first in each group, ties kept in index order."""
    from pyfe3d_amd.sweep import split_order
    bw = [65, 263, 137, 275, 257, 65, 256]
    sizes = [126, 900, 400, 900, 700, 126, 800]
    order, n_wide = split_order(bw, sizes)
    assert n_wide == 3
    assert order.dtype == np.int32
    assert order.tolist() == [1, 3, 4, 6, 2, 0, 5]
    order, n_wide = split_order([10, 20], [5, 9])
    assert n_wide == 0 and order.tolist() == [1, 0]
