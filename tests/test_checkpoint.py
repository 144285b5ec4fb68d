"""fluxfx-checkpoint-1 archives (backend/checkpoint.py:14-69): the loader's
validation runs before any device work, so malformed archives are rejected
on the CPU."""
import json

import numpy as np
import pytest

from fluxfx_amd.backend.checkpoint import FORMAT, load_checkpoint


def _archive(path, meta_update=None, arrays_update=None, n=4):
    meta = {"format": FORMAT, "settings": {}, "grid": {"shape": [n, n, n], "extent": [1.0, 1.0, 1.0]},
            "time": 0.5, "steps": 3, "last_dt": 0.01}
    meta.update(meta_update or {})
    cells = n ** 3
    arrays = {"density": np.zeros(cells, np.float32), "temperature": np.zeros(cells, np.float32),
              "pressure": np.zeros(cells, np.float32)}
    arrays.update({f"velocity{i}": np.zeros((n + 1) * n * n, np.float32) for i in range(3)})
    arrays.update(arrays_update or {})
    with open(path, "wb") as f:
        np.savez_compressed(f, metadata=json.dumps(meta), **arrays)
    return path


@pytest.mark.parametrize("meta,arrays,message", [
    ({"format": "fluxfx-checkpoint-0"}, None, "Unsupported checkpoint format"),
    ({"time": -1.0}, None, "Invalid checkpoint clock"),
    ({"steps": 1.5}, None, "Invalid checkpoint clock"),
    ({"last_dt": 0.5}, None, "timestep must be finite"),
    (None, {"density": np.zeros(3, np.float32)}, "Invalid field: density"),
    (None, {"temperature": np.full(64, np.nan, np.float32)}, "Invalid field: temperature"),
    (None, {"density": np.full(64, -1.0, np.float32)}, "Negative field: density"),
])
def test_malformed_archives_are_rejected(tmp_path, meta, arrays, message):
    path = _archive(tmp_path / "c.npz", meta, arrays)
    with pytest.raises(ValueError, match=message):
        load_checkpoint(path)


def test_unknown_settings_are_rejected(tmp_path):
    path = _archive(tmp_path / "c.npz", {"settings": {"no_such_field": 1}})
    with pytest.raises(TypeError):
        load_checkpoint(path)
