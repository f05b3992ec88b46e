"""Write the golden fixtures under tests/golden/ from the reference's own test data.

Run once, in a container that has the reference checkout:

    python tests/golden/make_golden.py /root/reference

Inputs (data files the reference's tests hold, ``tests/test_pipeline.py:18-130``):

* ``tests/fixtures/{orbium,orbium-scutium,aquarium}-test.yaml`` -- read with
  ``yaml.safe_load``; cells decoded by ``oracle/codec.py``;
* ``tests/fixtures/*_last_frame*.p`` -- numpy arrays stored as pickles.  They are
  NOT unpickled: ``static_ndarray`` walks the opcode stream with
  ``pickletools.genops`` (a disassembler) and evaluates only inert literals
  (ints, strings, bytes, tuples).  Global references stay strings and are never
  imported or called; the raw float32 payload is read with ``np.frombuffer``.

Outputs: ``<name>.json`` (v2 kernel list + world parameters),
``<name>_cells0.npy`` (initial state ``[1, C, H, W]``) and
``<name>_last_frame.npy`` (expected state after ``max_run_iter - 1`` steps).
"""
import json
import os
import pickletools
import sys

import numpy as np
import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import codec  # noqa: E402

CASES = {
    # name: (config yaml, golden frame)   -- tests/test_pipeline.py:19, 57, 95
    'orbium': ('orbium-test.yaml', 'orbium-test_last_frame.p'),
    'orbium_scutium': ('orbium-scutium-test.yaml', 'orbium-scutium-test_last_frame2.p'),
    'aquarium': ('aquarium-test.yaml', 'aquarium-test_last_frame.p'),
}

_ALLOWED = {
    'PROTO', 'FRAME', 'SHORT_BINUNICODE', 'BINUNICODE', 'MEMOIZE', 'BINPUT', 'LONG_BINPUT', 'BINGET',
    'LONG_BINGET', 'STACK_GLOBAL', 'GLOBAL', 'BININT1', 'BININT2', 'BININT', 'TUPLE1', 'TUPLE2', 'TUPLE3',
    'TUPLE', 'EMPTY_TUPLE', 'MARK', 'SHORT_BINBYTES', 'BINBYTES', 'REDUCE', 'NONE', 'NEWFALSE', 'NEWTRUE',
    'BUILD', 'STOP',
}


def static_ndarray(path):
    """Recover a little-endian float32 ndarray from a numpy pickle without unpickling it."""
    data = open(path, 'rb').read()
    stack, marks, memo = [], [], {}
    result = None
    for op, arg, _ in pickletools.genops(data):
        name = op.name
        if name not in _ALLOWED:
            raise ValueError(f'{path}: opcode {name} not accepted')
        if name in ('PROTO', 'FRAME'):
            continue
        if name in ('SHORT_BINUNICODE', 'BINUNICODE', 'BININT1', 'BININT2', 'BININT', 'SHORT_BINBYTES', 'BINBYTES'):
            stack.append(arg)
        elif name == 'NONE':
            stack.append(None)
        elif name == 'NEWFALSE':
            stack.append(False)
        elif name == 'NEWTRUE':
            stack.append(True)
        elif name == 'MEMOIZE':
            memo[len(memo)] = stack[-1]
        elif name in ('BINPUT', 'LONG_BINPUT'):
            memo[arg] = stack[-1]
        elif name in ('BINGET', 'LONG_BINGET'):
            stack.append(memo[arg])
        elif name == 'STACK_GLOBAL':
            nm = stack.pop()
            mod = stack.pop()
            stack.append(('global', mod, nm))
        elif name == 'GLOBAL':
            mod, nm = arg.split(' ', 1)
            stack.append(('global', mod, nm))
        elif name == 'MARK':
            marks.append(len(stack))
        elif name == 'TUPLE':
            m = marks.pop()
            t = tuple(stack[m:])
            del stack[m:]
            stack.append(t)
        elif name == 'EMPTY_TUPLE':
            stack.append(())
        elif name in ('TUPLE1', 'TUPLE2', 'TUPLE3'):
            n = int(name[-1])
            t = tuple(stack[-n:])
            del stack[-n:]
            stack.append(t)
        elif name == 'REDUCE':
            args = stack.pop()
            fn = stack.pop()
            stack.append(('call', fn, args))
        elif name == 'BUILD':
            state = stack.pop()
            obj = stack.pop()
            stack.append(('build', obj, state))
        elif name == 'STOP':
            result = stack.pop()
    # Expected shape: build(call(global numpy ... _reconstruct, ...), (1, shape, dtype, fortran, payload))
    tag, obj, state = result
    assert tag == 'build' and obj[0] == 'call' and obj[1][0] == 'global' and obj[1][2] == '_reconstruct'
    _version, shape, dtype, fortran, payload = state
    dtag, dcall, dstate = dtype
    assert dtag == 'build' and dcall[1][2] == 'dtype' and dcall[2][0] == 'f4' and dstate[1] == '<'
    assert fortran is False
    arr = np.frombuffer(payload, dtype='<f4')
    return arr.reshape(shape).copy()


def convert(ref_root, name, cfg_file, frame_file):
    fx = os.path.join(ref_root, 'tests', 'fixtures')
    with open(os.path.join(fx, cfg_file)) as fh:
        cfg = yaml.safe_load(fh)
    wp, rp = cfg['world_params'], cfg['render_params']
    nb_dims, C = wp['nb_dims'], wp['nb_channels']
    world_size = [2 ** rp['size_power2']] * nb_dims  # leniax/utils.py:53-55
    kernels = codec.update_config_v1_v2(cfg['kernels_params']['k'])
    raw = codec.load_raw_cells(cfg['run_params']['cells'], nb_dims, C)
    cells0 = codec.create_init_cells(world_size, C, [raw])
    frame = static_ndarray(os.path.join(fx, frame_file))
    meta = {
        'name': name,
        'world_size': world_size,
        'nb_channels': C,
        'R': wp['R'],
        'T': wp['T'],
        'weighted_average': wp.get('weighted_average', True),
        'get_state_fn_slug': wp.get('get_state_fn_slug', 'v1'),
        'max_run_iter': cfg['run_params']['max_run_iter'],
        'kernels_params': kernels,
        'source': {'config': f'tests/fixtures/{cfg_file}', 'frame': f'tests/fixtures/{frame_file}'},
    }
    with open(os.path.join(HERE, f'{name}.json'), 'w') as fh:
        json.dump(meta, fh, indent=1)
    np.save(os.path.join(HERE, f'{name}_cells0.npy'), cells0.astype(np.float32))
    np.save(os.path.join(HERE, f'{name}_last_frame.npy'), frame.astype(np.float32))
    print(name, 'cells0', cells0.shape, 'frame', frame.shape, 'raw', raw.shape)


if __name__ == '__main__':
    root = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
    for n, (c, fr) in CASES.items():
        convert(root, n, c, fr)
