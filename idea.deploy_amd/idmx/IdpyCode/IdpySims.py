"""Simulation snapshots: DumpSnapshot / ReadSnapshotData / CheckSnapshotData.

Follows idpy/IdpyCode/IdpySims.py:73-170. A snapshot holds one group named
after the simulation class with three sub-groups:

    <Class>/vars/<key>          numpy / builtin entries of sims_vars
                                (None is stored as an empty float32 array)
    <Class>/idpy_memory/<key>   D2H() of every device array (or of those
                                listed in sims_dump_idpy_memory)
    <Class>/custom_types/<key>  the type names of custom_types.Push(), bytes

The file is HDF5 when h5py can be imported (the reference's format, opened
in append mode, one group per class). Without h5py the same keys go into a
numpy .npz container written to exactly file_name (allow_pickle is never
used, so entries that would need pickling are skipped with a message, as the
reference skips non-numpy / non-builtin values). ReadSnapshotData tells the
two containers apart by their leading bytes, so callers never need to know
which one was written.
"""
import importlib.util
import os

import numpy as np

_HDF5_MAGIC = b'\x89HDF\r\n\x1a\n'


def have_h5py():
    return importlib.util.find_spec('h5py') is not None


def _is_hdf5(file_name):
    with open(file_name, 'rb') as f:
        return f.read(8) == _HDF5_MAGIC


def _storable(value):
    """The array an npz entry would hold, or None if it needs pickling."""
    try:
        a = np.asarray(value)
    except Exception:
        return None
    return None if a.dtype == object else a


def _collect(sims, custom_types):
    """(vars, idpy_memory, custom_types) entries of a snapshot, following
    the reference's selection rules (IdpySims.py:86-136)."""
    out_vars = {}
    if sims.sims_dump_vars_flag:
        for key, value in sims.sims_vars.items():
            if not ((len(sims.sims_dump_vars) == 0 and len(sims.sims_not_dump_vars) == 0)
                    or key not in sims.sims_not_dump_vars):
                continue
            kind = type(value).__module__.split('.')[0]
            if value is None:
                out_vars[key] = None
            elif kind in (np.__name__, 'builtins'):
                out_vars[key] = value
            else:
                print('Key: ', key, 'not builtin/numpy: not dumped!')
    out_mem = {}
    if sims.sims_dump_idpy_memory_flag:
        for key, arr in sims.sims_idpy_memory.items():
            if arr is None:
                raise Exception('Cannot dump ', key, ': not allocated')
            if len(sims.sims_dump_idpy_memory) == 0 or key in sims.sims_dump_idpy_memory:
                out_mem[key] = arr
    out_ct = {str(k): np.bytes_(v) for k, v in custom_types.Push().items()}
    return out_vars, out_mem, out_ct


def dump(sims, file_name, custom_types):
    if file_name is None:
        raise Exception('Parameter file_name must not be None')
    if custom_types is None:
        raise Exception('Parameter custom_types must not be None')
    cls = sims.__class__.__name__
    v, m, ct = _collect(sims, custom_types)
    if have_h5py():
        import h5py
        f = h5py.File(file_name, 'a')
        try:
            grp = f.create_group(cls)
            gv = grp.create_group('vars')
            for key, value in v.items():
                try:
                    gv.create_dataset(key, data=h5py.Empty('f') if value is None else value)
                except TypeError:
                    print('Key: ', key, 'not storable: not dumped!')
            gm = grp.create_group('idpy_memory')
            for key, arr in m.items():
                gm.create_dataset(key, data=arr.D2H())
            gc = grp.create_group('custom_types')
            for key, value in ct.items():
                gc.create_dataset(key, data=value)
        finally:
            f.close()
        return
    entries = {}
    if os.path.exists(file_name):
        if _is_hdf5(file_name):
            raise Exception('%s is an HDF5 file and h5py is not available' % file_name)
        with np.load(file_name, allow_pickle=False) as old:
            entries.update({k: old[k] for k in old.files})
    if any(k.split('/')[0] == cls for k in entries):
        raise ValueError('Unable to create group (name already exists): %s' % cls)
    for key, value in v.items():
        a = np.empty(0, dtype=np.float32) if value is None else _storable(value)
        if a is None:
            print('Key: ', key, 'not storable without pickling: not dumped!')
            continue
        entries['%s/vars/%s' % (cls, key)] = a
    for key, arr in m.items():
        entries['%s/idpy_memory/%s' % (cls, key)] = arr.D2H()
    for key, value in ct.items():
        entries['%s/custom_types/%s' % (cls, key)] = np.asarray(value)
    with open(file_name, 'wb') as f:
        np.savez(f, **entries)


def _open_check(sims, file_name, full_key):
    if file_name is None:
        raise Exception('Parameter file_name must not be None')
    if full_key is None:
        raise Exception('Parameter full_key must not be None')
    cls = sims.__class__.__name__
    if _is_hdf5(file_name):
        import h5py
        f = h5py.File(file_name, 'r')
        first = list(f.keys())[0]
        keys = f
    else:
        f = np.load(file_name, allow_pickle=False)
        first = f.files[0].split('/')[0] if f.files else None
        keys = set(f.files)
    if first != cls:
        print('File class: ', first)
        print('Present class: ', cls)
        f.close()
        raise Exception('The file you are reading has been created by another class!')
    return f, keys


def read(sims, file_name, full_key):
    f, keys = _open_check(sims, file_name, full_key)
    try:
        if full_key not in keys:
            raise Exception('Key', full_key, 'cannot be found in', file_name)
        return np.array(f.get(full_key))
    finally:
        f.close()


def check(sims, file_name, full_key):
    f, keys = _open_check(sims, file_name, full_key)
    f.close()
    return full_key in keys
