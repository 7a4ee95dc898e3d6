"""Persistent tuning results (the reference's SQLite tccg.db, tccg/sql_util.py:11-407).

The reference keys a measurement by the tensors' strings, their sizes, the CPU
model, the thread count, the float types and beta != 0
(tccg/sql_util.py:206-239) and stores every timed GETT candidate with its
measured and estimated GFLOPS (tccg/sql_util.py:97-154); on a later run it
re-emits the fastest candidate without timing (tccg/tccg.py:201-221).
Here the key is the (fused) contraction, the device architecture together
with a digest of the tile tables (tccg_amd.plan.tile_tables_digest), the
precision and beta != 0, and a GETT row records the gfx950 plan parameters
(tile id and its shape BM x BN x BK, m blocking, plan kind) with its measured
and estimated GFLOP/s.  A row is only reused when its tile id still names a
tile of the recorded shape.
"""
from __future__ import annotations

import os
import sqlite3
import time
from pathlib import Path
from typing import Optional, Tuple

SCHEMA = """
CREATE TABLE IF NOT EXISTS measurements (
  measurement_id INTEGER PRIMARY KEY AUTOINCREMENT,
  tensorA TEXT NOT NULL, tensorB TEXT NOT NULL, tensorC TEXT NOT NULL,
  device TEXT NOT NULL, prec TEXT NOT NULL, beta_nonzero INTEGER NOT NULL,
  created REAL NOT NULL,
  UNIQUE(tensorA, tensorB, tensorC, device, prec, beta_nonzero)
);
CREATE TABLE IF NOT EXISTS gett (
  gett_id INTEGER PRIMARY KEY AUTOINCREMENT,
  measurement_id INTEGER NOT NULL REFERENCES measurements(measurement_id),
  name TEXT NOT NULL, tile INTEGER NOT NULL, m_blocked INTEGER NOT NULL,
  gflops REAL NOT NULL, estimated_gflops REAL NOT NULL,
  kind INTEGER NOT NULL DEFAULT 0,
  bm INTEGER NOT NULL DEFAULT 0, bn INTEGER NOT NULL DEFAULT 0, bk INTEGER NOT NULL DEFAULT 0
);
"""


def default_path() -> Path:
    env = os.environ.get("TCCG_DB")
    if env:
        return Path(env)
    return Path(os.path.expanduser("~")) / ".cache" / "tccg_amd" / "tccg.db"


class TuningDB:
    def __init__(self, path: Optional[os.PathLike] = None):
        self.path = Path(path) if path else default_path()
        self.path.parent.mkdir(parents=True, exist_ok=True)
        self.con = sqlite3.connect(str(self.path))
        self.con.executescript(SCHEMA)
        cols = [r[1] for r in self.con.execute("PRAGMA table_info(gett)").fetchall()]
        # databases written before plan variants / tile shapes were recorded
        # (rows without a shape read as 0 x 0 x 0 and are never reused)
        for col in ("kind", "bm", "bn", "bk"):
            if col not in cols:
                self.con.execute("ALTER TABLE gett ADD COLUMN %s INTEGER NOT NULL DEFAULT 0" % col)
        self.con.commit()

    def close(self) -> None:
        self.con.close()

    def _key(self, contraction, device: str, prec: str, beta: float):
        return (str(contraction.A), str(contraction.B), str(contraction.C), device, prec, int(beta != 0))

    def measurement_id(self, contraction, device: str, prec: str, beta: float, create: bool = False) -> int:
        key = self._key(contraction, device, prec, beta)
        row = self.con.execute(
            "SELECT measurement_id FROM measurements WHERE tensorA=? AND tensorB=? AND tensorC=? AND device=? "
            "AND prec=? AND beta_nonzero=?", key).fetchone()
        if row:
            return int(row[0])
        if not create:
            return -1
        cur = self.con.execute(
            "INSERT INTO measurements(tensorA, tensorB, tensorC, device, prec, beta_nonzero, created) "
            "VALUES (?,?,?,?,?,?,?)", key + (time.time(),))
        self.con.commit()
        return int(cur.lastrowid)

    def insert_gett(self, mid: int, name: str, tile: int, m_blocked: bool, gflops: float, estimated: float,
                    kind: int = 0, shape: Tuple[int, int, int] = (0, 0, 0)) -> None:
        self.con.execute("INSERT INTO gett(measurement_id, name, tile, m_blocked, gflops, estimated_gflops, kind, "
                         "bm, bn, bk) VALUES (?,?,?,?,?,?,?,?,?,?)",
                         (mid, name, int(tile), int(bool(m_blocked)), float(gflops), float(estimated), int(kind),
                          int(shape[0]), int(shape[1]), int(shape[2])))
        self.con.commit()

    def fastest(self, contraction, device: str, prec: str,
                beta: float) -> Optional[Tuple[str, int, bool, int, float, Tuple[int, int, int]]]:
        """(name, tile, m_blocked, kind, gflops, (BM, BN, BK)) of the fastest
        stored candidate (kind: tccg_amd.plan.KIND_* bits)."""
        mid = self.measurement_id(contraction, device, prec, beta)
        if mid < 0:
            return None
        row = self.con.execute("SELECT name, tile, m_blocked, kind, gflops, bm, bn, bk FROM gett "
                               "WHERE measurement_id=? ORDER BY gflops DESC LIMIT 1", (mid,)).fetchone()
        if not row:
            return None
        return (str(row[0]), int(row[1]), bool(row[2]), int(row[3]), float(row[4]),
                (int(row[5]), int(row[6]), int(row[7])))
