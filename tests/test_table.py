"""The image-kernel table derived by mhs_amd/gen/mindlin_table.py is the
reference's, record for record: its canonical JSON hashes to the fingerprint
recorded from src/mhs/kernels/_image_table.py (tests/golden/make_golden.py)."""
import hashlib
import json

from conftest import GOLDEN, ROOT


def test_table_matches_reference_fingerprint():
    mine = json.loads((ROOT / "mhs_amd" / "data" / "mindlin_table.json").read_text())
    ref = json.loads((GOLDEN / "table_fingerprint.json").read_text())
    for key, want in ref.items():
        text = json.dumps(mine[key], separators=(",", ":"))
        assert len(mine[key]) == want["n"], key
        assert hashlib.sha256(text.encode()).hexdigest() == want["sha256"], key


def test_table_shape_facts():
    """Facts the reference asserts on load (image.py:65-92): record counts,
    odd R-family orders {3, 5, 7, 9}, rank <= 5, Q coefficients free of z."""
    t = json.loads((ROOT / "mhs_amd" / "data" / "mindlin_table.json").read_text())
    assert [len(t[k]) for k in ("r_dg", "r_ddg", "q_dg", "q_ddg")] == [163, 675, 195, 1134]
    orders = {r[6] for r in t["r_dg"]} | {r[7] for r in t["r_ddg"]}
    assert orders == {3, 5, 7, 9}
    assert max(sum(r[4:7]) for r in t["r_ddg"]) == 5
    assert all(term[0] == 0 for r in t["q_ddg"] for term in r[-1])
