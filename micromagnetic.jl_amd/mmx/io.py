"""State and table files in the reference's formats.

OVF 2.0 (src/ovf2.jl): header lines "# key: value", data "Text",
"Binary 4" (check value 1234567.0) or "Binary 8" (check value
123456789012345.0), little-endian, cells x fastest with (m_x, m_y, m_z) per
cell -- the order of Sim.spin.

Tables (src/fileio.jl:1-153): two comment lines of "%19s " names and units,
then one row per saved step, integers as "%18d " and floats as "%+18.12e ".
"""
import os

import numpy as np

OVF_CHECK4 = np.float32(1234567.0)
OVF_CHECK8 = np.float64(123456789012345.0)


def _jl_str(x):
    """Julia's string(x) for the header numbers (shortest round-trip digits)."""
    if isinstance(x, (int, np.integer)):
        return str(int(x))
    r = repr(float(x))
    if "e" in r:
        mant, exp = r.split("e")
        if "." not in mant:
            mant += ".0"
        return f"{mant}e{int(exp)}"
    return r


class OVF2:
    """The reference's OVF2 struct (src/ovf2.jl:1-12)."""

    def __init__(self, xnodes=1, ynodes=1, znodes=1, xstepsize=1e-9, ystepsize=1e-9, zstepsize=1e-9,
                 type="Float64", name="mag", data=None):
        self.xnodes, self.ynodes, self.znodes = int(xnodes), int(ynodes), int(znodes)
        self.xstepsize, self.ystepsize, self.zstepsize = float(xstepsize), float(ystepsize), float(zstepsize)
        self.type, self.name = type, name
        n = self.xnodes * self.ynodes * self.znodes
        self.data = np.zeros(3 * n) if data is None else np.asarray(data, dtype=np.float64).reshape(-1)


def _hdr(f, a, b):
    f.write(f"# {a}: {b}\n".encode())


def write_ovf(ovf, fname):
    """save_ovf(ovf, fname): src/ovf2.jl:71-83, :89-193."""
    if not fname.endswith(".ovf"):
        fname = fname + ".ovf"
    d = os.path.dirname(fname)
    if d:
        os.makedirs(d, exist_ok=True)
    nx, ny, nz = ovf.xnodes, ovf.ynodes, ovf.znodes
    dx, dy, dz = ovf.xstepsize, ovf.ystepsize, ovf.zstepsize
    with open(fname, "wb") as f:
        f.write(b"# OOMMF OVF 2.0\n")
        _hdr(f, "Segment count", "1")
        _hdr(f, "Begin", "Segment")
        _hdr(f, "Begin", "Header")
        _hdr(f, "Title", ovf.name)
        _hdr(f, "meshtype", "rectangular")
        _hdr(f, "meshunit", "m")
        for k in ("xmin", "ymin", "zmin"):
            _hdr(f, k, 0)
        _hdr(f, "xmax", _jl_str(nx * dx))
        _hdr(f, "ymax", _jl_str(ny * dy))
        _hdr(f, "zmax", _jl_str(nz * dz))
        _hdr(f, "valuedim", 3)
        _hdr(f, "valuelabels", "m_x m_y m_z")
        _hdr(f, "valueunits", "1 1 1")
        _hdr(f, "xbase", _jl_str(dx / 2))
        _hdr(f, "ybase", _jl_str(dy / 2))
        _hdr(f, "zbase", _jl_str(dz / 2))
        _hdr(f, "xnodes", nx)
        _hdr(f, "ynodes", ny)
        _hdr(f, "znodes", nz)
        _hdr(f, "xstepsize", _jl_str(dx))
        _hdr(f, "ystepsize", _jl_str(dy))
        _hdr(f, "zstepsize", _jl_str(dz))
        _hdr(f, "End", "Header")
        if ovf.type in ("String", str):
            _hdr(f, "Begin", "Data Text")
            b = ovf.data.reshape(-1, 3)
            for v in b:
                f.write(f"{_jl_str(v[0])} {_jl_str(v[1])} {_jl_str(v[2])}\n".encode())
            _hdr(f, "End", "Data Text")
        elif ovf.type in ("Float32", np.float32):
            _hdr(f, "Begin", "Data Binary 4")
            f.write(OVF_CHECK4.astype("<f4").tobytes())
            f.write(ovf.data.astype("<f4").tobytes())
            f.write(b"\n")
            _hdr(f, "End", "Data Binary 4")
        else:
            _hdr(f, "Begin", "Data Binary 8")
            f.write(OVF_CHECK8.astype("<f8").tobytes())
            f.write(ovf.data.astype("<f8").tobytes())
            f.write(b"\n")
            _hdr(f, "End", "Data Binary 8")
        _hdr(f, "End", "Segment")
    return fname


def read_ovf_file(fname):
    """read_ovf(fname): src/ovf2.jl:268-313 -> OVF2 (data as float64)."""
    if not fname.endswith(".ovf"):
        fname = fname + ".ovf"
    ovf = OVF2()
    with open(fname, "rb") as f:
        if not f.readline().decode(errors="replace").startswith("# OOMMF OVF 2"):
            raise ValueError("Input is not a OVF2 file.")
        while True:
            raw = f.readline()
            if not raw:
                break
            line = raw.decode(errors="replace").rstrip("\n")
            key = line[2:].split(":", 1)[0] if line.startswith("# ") else ""
            val = line.split(":", 1)[1].strip() if ":" in line else ""
            if key == "Title":
                ovf.name = line[9:]
            elif key in ("xnodes", "ynodes", "znodes"):
                setattr(ovf, key, int(val))
            elif key in ("xstepsize", "ystepsize", "zstepsize"):
                setattr(ovf, key, float(val))
            elif line.startswith("# Begin: Data"):
                kind = line[14:]
                n = 3 * ovf.xnodes * ovf.ynodes * ovf.znodes
                if kind == "Binary 8":
                    ovf.type = "Float64"
                    chk = np.frombuffer(f.read(8), "<f8")[0]
                    ovf.data = np.frombuffer(f.read(8 * n), "<f8").astype(np.float64) if chk == OVF_CHECK8 \
                        else np.zeros(n)
                elif kind == "Binary 4":
                    ovf.type = "Float32"
                    chk = np.frombuffer(f.read(4), "<f4")[0]
                    ovf.data = np.frombuffer(f.read(4 * n), "<f4").astype(np.float64) if chk == OVF_CHECK4 \
                        else np.zeros(n)
                elif kind.lower() == "text":
                    ovf.type = "String"
                    vals = []
                    for raw2 in f:
                        l2 = raw2.decode()
                        if l2.startswith("#"):
                            break
                        vals.extend(float(x) for x in l2.split())
                    ovf.data = np.zeros(n)
                    ovf.data[:len(vals)] = vals
                else:
                    raise ValueError("Data format error!")
                break
    return ovf


def save_ovf(sim, fname, type="Float64"):
    """save_ovf(sim, fname; type): src/ovf2.jl:46-64."""
    m = sim.mesh
    ovf = OVF2(m.nx, m.ny, m.nz, m.dx, m.dy, m.dz, type=type, name=getattr(sim, "name", "dyn"), data=sim.spin)
    return write_ovf(ovf, fname)


def read_ovf(sim, fname):
    """read_ovf(sim, fname): src/ovf2.jl:200-208 -- the state becomes the file's m."""
    from . import init_m0

    ovf = read_ovf_file(fname)
    if ovf.xnodes * ovf.ynodes * ovf.znodes != sim.n_total:
        raise ValueError("The ovf does not match the sim.mesh")
    init_m0(sim, ovf.data, norm=False)
    return ovf


def mag2ovf(m, nx, ny, nz, dx=1e-9, dy=1e-9, dz=1e-9, fname="mag", type="Float64"):
    """mag2ovf: src/ovf2.jl:330-343."""
    return write_ovf(OVF2(nx, ny, nz, dx, dy, dz, type=type, name=fname, data=m), fname)


# ---- tables --------------------------------------------------------------------
def _fmt(x):
    if isinstance(x, tuple):
        return "".join(_fmt(v) for v in x)
    if isinstance(x, str):
        return "%19s " % x
    if isinstance(x, (int, np.integer)):
        return "%18d " % x
    return "%+18.12e " % x


class SaverItem:
    def __init__(self, name, unit, result):
        self.name, self.unit, self.result = name, unit, result


class DataSaver:
    """DataSaver + write_data: src/fileio.jl:39-104.  Default items as
    create_saver + the columns the add_* calls register, in insertion order:
    step, time, E_total, m_x, m_y, m_z, then per term its field columns
    (Zeeman), E_<term> (not for the torques) and T_<term> (thermal)."""

    def __init__(self, name, items=None):
        self.name = name
        self.header_saved = False
        self.items = list(items or [])
        self.t = 0.0
        self.nsteps = 0

    @classmethod
    def for_sim(cls, sim, name=None):
        from . import average_m, compute_system_energy

        s = cls(name or f"{getattr(sim, 'name', 'dyn')}_llg.txt")
        state = {}

        def energies(o):  # compute_system_energy(sim, sim.spin, t_end), sim.jl:555,569
            if state.get("at") != s.nsteps:
                state["E"] = compute_system_energy(o, t=s.t)
                state["at"] = s.nsteps
            return state["E"]

        s.items.append(SaverItem("step", "<unitless>", lambda o: s.nsteps))
        s.items.append(SaverItem("time", "<s>", lambda o: s.t))
        s.items.append(SaverItem("E_total", "<J>", lambda o: float(np.sum(energies(o)))))
        s.items.append(SaverItem(("m_x", "m_y", "m_z"), ("<unitless>", "<unitless>", "<unitless>"),
                                 lambda o: tuple(float(v) for v in average_m(o))))
        for q, term in enumerate(sim.interactions):
            for names, units, fn in getattr(term, "items_before", ()):
                s.items.append(SaverItem(names, units, lambda o, fn=fn: fn(o, s)))
            unit = getattr(term, "energy_unit", "<J>")
            if unit is not None:
                s.items.append(SaverItem(f"E_{term.name}", unit, lambda o, q=q: float(energies(o)[q])))
            for names, units, fn in getattr(term, "items_after", ()):
                s.items.append(SaverItem(names, units, lambda o, fn=fn: fn(o, s)))
        return s

    def write(self, sim):
        mode = "a" if self.header_saved else "w"
        with open(self.name, mode) as f:
            if not self.header_saved:
                f.write("#" + "".join(_fmt(it.name) for it in self.items) + "\n")
                f.write("#" + "".join(_fmt(it.unit) for it in self.items) + "\n")
                self.header_saved = True
            f.write("".join(_fmt(it.result(sim)) for it in self.items) + "\n")


def read_table(filename):
    """read_table: src/fileio.jl:125-153 -> (data dict, units dict)."""
    with open(filename) as f:
        lines = f.read().splitlines()
    start = next(i for i, l in enumerate(lines) if not l.startswith("#"))
    headers = lines[start - 2].replace("#", "").split()
    units = lines[start - 1].replace("#", "").replace("<", "").replace(">", "").split()
    units = ["unitless" if (u == "1" or not u) else u for u in units]
    data = {h: [] for h in headers}
    for l in lines[start:]:
        vals = l.split()
        for i, h in enumerate(headers):
            data[h].append(float(vals[i]))
    return {h: np.array(v) for h, v in data.items()}, dict(zip(headers, units))
