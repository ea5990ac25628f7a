"""The .mtz mesh tarball reader (MeshTarball, src/bem/mesh_tarball.cpp:
118-270, as Mesh::init loads it, src/bem/mesh.cpp:88-151): archives are built
here with Python's tarfile from a mesh written by the library, and must read
back to the same surface, charges, centre and initial f/h."""
import gzip
import io
import tarfile

import numpy as np
import pytest

from beep3_amd._lib import BeepError
from beep3_amd.scene import Mesh

DEF = """<?xml version="1.0" ?>
<!-- mesh tarball definition -->
<contents>
  <mesh>{mesh}</mesh>
  <xyzqr>charges.xyzqr</xyzqr>
  <centre>centre.txt</centre>
  <energies>energies.txt</energies>
  {extra}
</contents>
"""


def _members(tmp_path, mesh_name="surface.off", extra="", fh=None):
    m = Mesh.icosphere(2, 3.0)
    off = tmp_path / "surface.off"
    m.write_off(str(off))
    xyzqr = "# x y z q r\n0.5 0.0 0.1 1.0 1.5\n-0.4 0.3 0.0 -0.5 1.2\n"
    files = {
        "definition.xml": DEF.format(mesh=mesh_name, extra=extra),
        mesh_name: off.read_text(),
        "charges.xyzqr": xyzqr,
        "centre.txt": "0.25 -0.5 1.0\n",
        "energies.txt": "0\n",
    }
    if fh is not None:
        files["fh.txt"] = "".join(f"{float(a)!r} {float(b)!r}\n" for a, b in zip(*fh))
    return m, files


def _tar(path, files, mode="w:gz", prefix="./"):
    with tarfile.open(path, mode) as t:
        for name, text in files.items():
            data = text.encode()
            ti = tarfile.TarInfo(prefix + name)
            ti.size = len(data)
            t.addfile(ti, io.BytesIO(data))


@pytest.mark.parametrize("mode", ["w:gz", "w", "w:gz-pax"])
def test_mtz_round_trip(tmp_path, mode):
    n_v = Mesh.icosphere(2, 3.0).info()["n_vertices"]
    rng = np.random.default_rng(3)
    fh = (rng.normal(size=n_v), rng.normal(size=n_v))
    ref, files = _members(tmp_path, extra="<fh>fh.txt</fh><pdb>x.pdb</pdb>", fh=fh)
    p = tmp_path / "mol.mtz"
    if mode == "w:gz-pax":
        # a long member name (pax path record) for the surface
        long = "d" * 120 + "/surface.off"
        files[long] = files.pop("surface.off")
        files["definition.xml"] = files["definition.xml"].replace("surface.off", long)
        with tarfile.open(p, "w:gz", format=tarfile.PAX_FORMAT) as t:
            for name, text in files.items():
                data = text.encode()
                ti = tarfile.TarInfo(name)
                ti.size = len(data)
                t.addfile(ti, io.BytesIO(data))
    else:
        _tar(p, files, mode)
    m = Mesh.read_mtz(str(p))
    xyz0, tri0 = ref.vertices()
    xyz1, tri1 = m.vertices()
    assert np.array_equal(tri0, tri1)
    assert np.allclose(xyz0, xyz1, rtol=0, atol=1e-15)
    i = m.info()
    assert i["n_charges"] == 2
    assert np.array_equal(i["centre"], [0.25, -0.5, 1.0])  # the centre member, not the charge centroid
    f, h = m.initial_fh()
    assert np.array_equal(f, fh[0]) and np.array_equal(h, fh[1])
    # the centre stays when charges are replaced
    m.set_charges([[1.0, 1.0, 1.0, 1.0]])
    assert np.array_equal(m.info()["centre"], [0.25, -0.5, 1.0])


def test_mtz_gts_member_and_no_fh(tmp_path):
    ref, files = _members(tmp_path)
    # the same surface as GTS: vertices, edges (1-based), faces as edge triples
    xyz, tri = ref.vertices()
    edges, faces = {}, []
    for a, b, c in tri:
        fe = []
        for u, v in ((a, b), (b, c), (c, a)):
            key = (min(u, v), max(u, v))
            if key not in edges:
                edges[key] = len(edges) + 1
            fe.append(edges[key])
        faces.append(fe)
    lines = [f"{len(xyz)} {len(edges)} {len(faces)}"]
    lines += [f"{float(x)!r} {float(y)!r} {float(z)!r}" for x, y, z in xyz]
    lines += [f"{u + 1} {v + 1}" for (u, v), _ in sorted(edges.items(), key=lambda kv: kv[1])]
    lines += [" ".join(map(str, fe)) for fe in faces]
    files.pop("surface.off")
    files["surface.gts"] = "\n".join(lines) + "\n"
    files["definition.xml"] = files["definition.xml"].replace("surface.off", "surface.gts")
    p = tmp_path / "mol.mtz"
    _tar(p, files, prefix="")
    m = Mesh.read_mtz(str(p))
    # the GTS reader takes each face's corners from consecutive edge pairs
    # (gts_utils.h:99-120): the same triangles, cyclically rotated
    def canon(t):
        k = np.argmin(t, axis=1)
        return np.array([np.roll(r, -i) for r, i in zip(t, k)])
    assert np.array_equal(canon(m.vertices()[1]), canon(tri))
    assert np.allclose(m.vertices()[0], xyz, rtol=0, atol=1e-15)
    assert m.initial_fh() is None


def test_mtz_errors(tmp_path):
    ref, files = _members(tmp_path)
    bad = dict(files)
    bad["definition.xml"] = files["definition.xml"].replace("<centre>centre.txt</centre>", "")
    _tar(tmp_path / "a.mtz", bad)
    with pytest.raises(BeepError, match="centre"):
        Mesh.read_mtz(str(tmp_path / "a.mtz"))
    bad = dict(files)
    bad["definition.xml"] = files["definition.xml"].replace("</contents>", "<mesh>x.off</mesh></contents>")
    _tar(tmp_path / "b.mtz", bad)
    with pytest.raises(BeepError, match="multiple"):
        Mesh.read_mtz(str(tmp_path / "b.mtz"))
    bad = dict(files)
    bad.pop("charges.xyzqr")
    _tar(tmp_path / "c.mtz", bad)
    with pytest.raises(BeepError, match="missing"):
        Mesh.read_mtz(str(tmp_path / "c.mtz"))
    bad = dict(files)
    bad["definition.xml"] = files["definition.xml"].replace("contents>", "stuff>")
    _tar(tmp_path / "d.mtz", bad)
    with pytest.raises(BeepError, match="contents"):
        Mesh.read_mtz(str(tmp_path / "d.mtz"))
    _tar(tmp_path / "e.mtz", files, mode="w:bz2")
    with pytest.raises(BeepError, match="bzip2"):
        Mesh.read_mtz(str(tmp_path / "e.mtz"))
    (tmp_path / "f.mtz").write_bytes(gzip.compress(b"not a tar archive" * 64))
    with pytest.raises(BeepError, match="tar"):
        Mesh.read_mtz(str(tmp_path / "f.mtz"))
