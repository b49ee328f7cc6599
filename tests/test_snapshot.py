"""Snapshot layout and selection rules of DumpSnapshot / ReadSnapshotData
(idpy/IdpyCode/IdpySims.py:73-176), on host-backed arrays (no GPU)."""
import numpy as np
import pytest


class _HostArray:
    def __init__(self, a):
        self.a = np.asarray(a)

    def D2H(self):
        return self.a.copy()


def _sim(cls_name='ShanChenMultiPhase'):
    from idmx.LBM._base import IdpySims
    cls = type(cls_name, (IdpySims,), {})
    s = cls()
    s.sims_vars.update({'V': 16, 'tau': 0.8, 'dim_sizes': np.array([4, 4]),
                        'label': 'droplet', 'empty': None, 'mixed': [1, 'a', None],
                        'f': (lambda x: x)})
    rng = np.random.default_rng(3)
    s.sims_idpy_memory.update({'pop': _HostArray(rng.random(9 * 16)),
                               'n': _HostArray(rng.random(16))})
    return s


class _CT:
    def Push(self):
        return {'PopType': 'double', 'NType': 'double'}


def test_dump_read_round_trip(tmp_path):
    s = _sim()
    fn = str(tmp_path / 'snap.hdf5')
    s.DumpSnapshot(file_name=fn, custom_types=_CT())
    c = 'ShanChenMultiPhase'
    assert np.array_equal(s.ReadSnapshotData(fn, c + '/idpy_memory/pop'),
                          s.sims_idpy_memory['pop'].a)
    assert np.array_equal(s.ReadSnapshotData(fn, c + '/idpy_memory/n'),
                          s.sims_idpy_memory['n'].a)
    assert s.ReadSnapshotData(fn, c + '/vars/V') == 16
    assert s.ReadSnapshotData(fn, c + '/vars/tau') == 0.8
    assert np.array_equal(s.ReadSnapshotData(fn, c + '/vars/dim_sizes'), [4, 4])
    assert str(s.ReadSnapshotData(fn, c + '/vars/label')) == 'droplet'
    assert s.ReadSnapshotData(fn, c + '/vars/empty').size == 0
    assert s.ReadSnapshotData(fn, c + '/custom_types/PopType') == b'double'
    # a function is neither numpy nor builtin; a mixed list needs pickling
    assert not s.CheckSnapshotData(fn, c + '/vars/f')
    assert s.CheckSnapshotData(fn, c + '/vars/V')
    with pytest.raises(Exception):
        s.ReadSnapshotData(fn, c + '/vars/nope')


def test_selection_rules_and_class_check(tmp_path):
    s = _sim()
    s.sims_not_dump_vars = ['tau']
    s.sims_dump_idpy_memory = ['n']
    fn = str(tmp_path / 'snap.hdf5')
    s.DumpSnapshot(file_name=fn, custom_types=_CT())
    c = 'ShanChenMultiPhase'
    assert not s.CheckSnapshotData(fn, c + '/vars/tau')
    assert s.CheckSnapshotData(fn, c + '/vars/V')
    assert not s.CheckSnapshotData(fn, c + '/idpy_memory/pop')
    assert s.CheckSnapshotData(fn, c + '/idpy_memory/n')
    # a second dump of the same class into the same file is refused
    with pytest.raises(Exception):
        s.DumpSnapshot(file_name=fn, custom_types=_CT())
    other = _sim('RootLB')
    with pytest.raises(Exception):
        other.ReadSnapshotData(fn, 'RootLB/vars/V')
    with pytest.raises(Exception):
        s.DumpSnapshot(file_name=None, custom_types=_CT())
    s.sims_idpy_memory['u'] = None
    with pytest.raises(Exception):
        s.DumpSnapshot(file_name=str(tmp_path / 'b.hdf5'), custom_types=_CT())
