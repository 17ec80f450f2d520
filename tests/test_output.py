"""VTU and CSV output of the mirror (src/hex_mesher.py:1446-1499,
src/topopt_run_benchmarks.py:218-250), read back with plain XML parsing."""
import base64
import xml.etree.ElementTree as ET

import numpy as np

from pyto_amd.hex_mesher import HexMesher
from pyto_amd.topopt_common import write_summary_row


def _read_array(el):
    raw = base64.b64decode(el.text)
    n = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
    dt = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}[el.get("type")]
    return np.frombuffer(raw[8:8 + n], dtype=dt)


def test_vtu_roundtrip(tmp_path):
    m = HexMesher()
    m.grid_mesh((3, 2, 2), (1.0, 0.5, 2.0))
    rng = np.random.default_rng(1)
    field = rng.uniform(size=m.num_elems)
    m.setPseudoDensity(field)
    path = str(tmp_path / "out" / "mesh.vtu")
    m.export_vtu_mesh(field, mask_low_pseudodensity=True, density_field="rho", file_name=path)
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    keep = field > 0.5
    assert int(piece.get("NumberOfPoints")) == m.num_nodes
    assert int(piece.get("NumberOfCells")) == int(keep.sum())
    arrays = {a.get("Name"): _read_array(a) for a in root.iter("DataArray")}
    assert np.array_equal(arrays["connectivity"].reshape(-1, 8), m.elemArray[keep])
    assert np.array_equal(arrays["offsets"], 8 * np.arange(1, keep.sum() + 1))
    assert np.all(arrays["types"] == 12)
    assert np.allclose(arrays["Points"].reshape(-1, 3), m.node_xyz)
    # point value = mean over the kept cells that share the point (brute force)
    want = np.zeros(m.num_nodes)
    for n in range(m.num_nodes):
        vals = [field[e] for e in np.nonzero(keep)[0] if n in m.elemArray[e]]
        want[n] = np.mean(vals) if vals else 0.0
    assert np.allclose(arrays["rho"], want, rtol=0, atol=1e-15)


def test_summary_csv_update_and_append(tmp_path):
    import pandas as pd
    path = str(tmp_path / "r" / "OC_summary.csv")
    write_summary_row(path, {"name": "a", "objective": 1.5, "success": True})
    write_summary_row(path, {"name": "b", "objective": 2.5, "success": False})
    write_summary_row(path, {"name": "a", "objective": 0.5, "success": True})
    df = pd.read_csv(path)
    assert list(df["name"]) == ["a", "b"]
    assert list(df["objective"]) == [0.5, 2.5]
