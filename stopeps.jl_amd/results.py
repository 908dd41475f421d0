"""Result and checkpoint files: ``Results.save_result`` / ``load_result``
(src/io/results.jl:8-80) and ``save_checkpoint`` / ``load_checkpoint``
(src/io/run_output.jl:1-39) with the reference's groups, keys, schema version,
checkpoint format string, atomic publication and checks.

The container differs: the reference writes JLD2 (HDF5), for which no library
exists in this environment, so a record is one ``.npz`` file (a zip archive,
read here without unpickling) holding a JSON document of the groups
(``schema_version``, ``created_at``, ``method``, ``model``, ``parameters``,
``observables``, ``diagnostics``, ``metadata``) and the numeric arrays it
refers to. Files of the reference cannot be read or written by this module.
"""
import datetime
import json
import math
import os
import tempfile

import numpy as np

RESULT_SCHEMA_VERSION = "2.7.0"
CHECKPOINT_FORMAT = "stopeps-tensorkit-thermal-1.0.0"
ACCEPTED_SCHEMAS = ("2.0.0", "2.1.0", "2.2.0", "2.3.0", "2.4.0", "2.5.0", "2.6.0", RESULT_SCHEMA_VERSION)
_INLINE = 64  # arrays with at most this many entries stay in the JSON document


def _encode(obj, arrays):
    if isinstance(obj, np.ndarray):
        if obj.size <= _INLINE and obj.dtype.kind in "biuf":
            return {"__ndarray__": [_encode(v, arrays) for v in obj.reshape(-1).tolist()],
                    "dtype": obj.dtype.str, "shape": list(obj.shape)}
        key = f"a{len(arrays)}"
        arrays[key] = np.ascontiguousarray(obj)
        return {"__array__": key}
    if isinstance(obj, dict):
        return {"__dict__": [[_encode(k, arrays), _encode(v, arrays)] for k, v in obj.items()]}
    if isinstance(obj, tuple):
        return {"__tuple__": [_encode(v, arrays) for v in obj]}
    if isinstance(obj, list):
        return [_encode(v, arrays) for v in obj]
    if isinstance(obj, (np.floating, float)):
        v = float(obj)
        return {"__float__": repr(v)} if not math.isfinite(v) else v
    if isinstance(obj, (np.integer,)):
        return int(obj)
    if isinstance(obj, np.bool_):
        return bool(obj)
    if obj is None or isinstance(obj, (bool, int, str)):
        return obj
    raise TypeError(f"cannot store a {type(obj).__name__} in a result file")


def _decode(obj, arrays):
    if isinstance(obj, list):
        return [_decode(v, arrays) for v in obj]
    if isinstance(obj, dict):
        if "__array__" in obj:
            return arrays[obj["__array__"]]
        if "__ndarray__" in obj:
            return np.array([_decode(v, arrays) for v in obj["__ndarray__"]],
                            dtype=np.dtype(obj["dtype"])).reshape(obj["shape"])
        if "__dict__" in obj:
            return {_decode(k, arrays): _decode(v, arrays) for k, v in obj["__dict__"]}
        if "__tuple__" in obj:
            return tuple(_decode(v, arrays) for v in obj["__tuple__"])
        if "__float__" in obj:
            return float(obj["__float__"])
    return obj


def save_result(path, method, model, parameters=None, observables=None, diagnostics=None, metadata=None,
                overwrite=True):
    """``Results.save_result`` (results.jl:28-74): write one versioned record
    to a temporary file in the destination directory and publish it
    atomically (a rename; with ``overwrite=False`` a hard link that fails when
    the destination exists). Returns the absolute destination."""
    if not str(path).lower().endswith(".npz"):
        raise ValueError("result path must use the .npz extension")
    destination = os.path.abspath(path)
    if not overwrite and (os.path.lexists(destination)):
        raise FileExistsError(f"output exists: {destination}; use overwrite=True to replace it")
    directory = os.path.dirname(destination)
    os.makedirs(directory, exist_ok=True)
    arrays = {}
    doc = {"schema_version": RESULT_SCHEMA_VERSION,
           "created_at": datetime.datetime.now(datetime.timezone.utc).isoformat(),
           "method": str(method), "model": _encode(model, arrays),
           "parameters": _encode({str(k): v for k, v in (parameters or {}).items()}, arrays),
           "observables": _encode(observables if observables is not None else {}, arrays),
           "diagnostics": _encode(diagnostics if diagnostics is not None else {}, arrays),
           "metadata": _encode({"runtime": "stopeps.jl_amd (Python host, HIP device)",
                                **{str(k): v for k, v in (metadata or {}).items()}}, arrays)}
    fd, tmp = tempfile.mkstemp(dir=directory, suffix=".npz")
    os.close(fd)
    try:
        arrays["__document__"] = np.frombuffer(json.dumps(doc).encode(), dtype=np.uint8)
        with open(tmp, "wb") as fh:
            np.savez_compressed(fh, **arrays)
        if overwrite:
            os.replace(tmp, destination)
        else:
            os.link(tmp, destination)
            os.unlink(tmp)
    except BaseException:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise
    return destination


def load_result(path):
    """``Results.load_result`` (results.jl:76-80): the record as a dict."""
    with np.load(path, allow_pickle=False) as f:
        arrays = {k: f[k] for k in f.files}
    doc = json.loads(bytes(arrays.pop("__document__")).decode())
    out = {}
    for k, v in doc.items():
        out[k] = v if k in ("schema_version", "created_at", "method") else _decode(v, arrays)
    return out


def state_sites(state):
    """Site tensors of a host ``FiniteState`` as stored in a checkpoint:
    per row-major site, the dense array and its six parity vectors."""
    return [{"data": state[xy].data, "parity": [np.asarray(p) for p in state[xy].parity]}
            for xy in state.lattice.site_order()]


def save_checkpoint(path, cp, overwrite=False):
    """``save_checkpoint`` (run_output.jl:13-21): the beta0 PEPO, its beta and
    log_amplitude and the purification rows in ``observables``, the
    preparation parameters with the checkpoint format, the history, the run
    id. A resident PEPO is read back from the device first (with its live bond
    sectors)."""
    from .sampling import DeviceState
    state = cp.state
    live = None
    if isinstance(state, DeviceState):
        live = state.live()
        state = state.download()
    params = dict(cp.preparation)
    params["checkpoint_format"] = CHECKPOINT_FORMAT
    obs = {"sites": state_sites(state), "beta0": state.beta, "log_amplitude": state.log_amplitude,
           "lattice": (state.lattice.Lx, state.lattice.Ly), "purification": list(cp.purification)}
    if live is not None:
        obs["live"] = live
    return save_result(path, method="stoPEPS_checkpoint", model=cp.preparation["model"], parameters=params,
                       observables=obs, diagnostics={"history": list(cp.history)},
                       metadata={"run_id": cp.run_id, "complete": True}, overwrite=overwrite)


def load_checkpoint(path, c, ctx=None):
    """``load_checkpoint`` (run_output.jl:28-39): check the schema, the
    checkpoint format and the model, lattice and beta0 against the run
    configuration, and rebuild the ``ThermalCheckpoint`` (its PEPO uploaded to
    ``ctx`` with its live bond sectors when a context is given, else a host
    ``FiniteState``)."""
    from .config import parameter_dict
    from .driver import ThermalCheckpoint
    from .particles import same_beta
    from .sampling import DeviceState
    from .states import FiniteState, Lattice, Site
    d = load_result(path)
    p = d["parameters"]
    if not (d["schema_version"] in ACCEPTED_SCHEMAS and p.get("checkpoint_format") == CHECKPOINT_FORMAT):
        raise ValueError("unsupported checkpoint format")
    o = d["observables"]
    current = parameter_dict(c)
    if not (p["model"] == current["model"] and p["lattice"] == current["lattice"] and same_beta(o["beta0"], c.beta0)):
        raise ValueError("checkpoint model, lattice or beta0 mismatch")
    Lx, Ly = c.lattice.Lx, c.lattice.Ly
    if tuple(o["lattice"]) != (Lx, Ly):
        raise ValueError("checkpoint model, lattice or beta0 mismatch")
    sites = [[None] * Ly for _ in range(Lx)]
    for k, s in enumerate(o["sites"]):
        sites[k % Lx][k // Lx] = Site(s["data"], s["parity"])
    state = FiniteState(Lattice(Lx, Ly), sites, "pepo", beta=o["beta0"], log_amplitude=o["log_amplitude"])
    symmetric = bool(np.any(state[(0, 0)].parity[2]))
    if symmetric != c.model.symmetric or state[(0, 0)].data.shape[2] != 2:
        raise ValueError("checkpoint physical space mismatch")
    if ctx is not None:
        state = DeviceState(ctx, state, live=o.get("live"))
    return ThermalCheckpoint(state, p, list(o["purification"]), list(d["diagnostics"]["history"]),
                             str(d["metadata"]["run_id"]))
