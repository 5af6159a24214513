"""VTK ImageData (.vti) output, mirroring exportSdfToVTI
(src/DataExport/ExportToVTI.jl:22-67), and the projected-node filter
SelectProjectedNodes (src/DataExport/DataPostProcess.jl:21-58).

The file carries the same content as the reference's WriteVTK output: a
point field named ``value_label`` on an ImageData grid with origin
AABB_min, spacing cell_size (or cell_size/smooth) and (N+1) (or N*smooth+1)
points per axis, x fastest.  Data are appended raw and zlib-compressed with
a UInt64 block header (WriteVTK's defaults).
"""
import zlib

import numpy as np

_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32"}


def SelectProjectedNodes(mesh, grid, xp, points):
    """SelectProjectedNodes(mesh, grid, xp, points) -- DataPostProcess.jl:21-58.

    Keeps the points whose projection is non-zero (sum |xp[:, i]| > 1e-10)
    and returns (X, Xp, mean_PD, max_PD): X and Xp as 3 x count arrays (the
    reference's vectors of 3-vectors, same order), and the mean and maximum
    of |X - Xp|; 0.0 and 0.0 when nothing is kept.  Host-side filtering of
    arrays the caller already holds, like the reference's loop."""
    ngp = int(grid.ngp)
    xp = np.asarray(xp, dtype=np.float64).reshape(3, -1)
    points = np.asarray(points, dtype=np.float64).reshape(3, -1)
    if xp.shape[1] != ngp or points.shape[1] != ngp:
        raise ValueError(f"xp and points must be 3 x {ngp}")
    keep = np.abs(xp).sum(axis=0) > 1.0e-10
    X = np.ascontiguousarray(points[:, keep])
    Xp = np.ascontiguousarray(xp[:, keep])
    if X.shape[1] == 0:
        return X, Xp, 0.0, 0.0
    pd = np.sqrt(((X - Xp) ** 2).sum(axis=0))
    return X, Xp, float(pd.mean()), float(pd.max())


def exportSdfToVTI(filename, grid, values, value_label, smooth=None, block=32768 * 8):
    dims = (np.asarray(grid.N, np.int64) + 1) if smooth is None else \
        (np.asarray(grid.N, np.int64) * int(smooth) + 1)
    h = float(grid.cell_size) if smooth is None else float(grid.cell_size) / int(smooth)
    v = np.ascontiguousarray(values)
    if v.dtype not in _TYPES:
        v = v.astype(np.float64)
    if v.ndim != 1 or v.size != int(np.prod(dims)):
        # ExportToVTI.jl:43-47
        raise ValueError(f"Values vector length ({v.size}) doesn't match grid dimensions "
                         f"({int(np.prod(dims))}).")
    if not filename.endswith(".vti"):
        filename = filename + ".vti"
    raw = v.tobytes()
    chunks = [raw[i:i + block] for i in range(0, len(raw), block)] or [b""]
    comp = [zlib.compress(c) for c in chunks]
    last = len(chunks[-1])
    header = np.array([len(chunks), block, last] + [len(c) for c in comp], dtype="<u8").tobytes()
    ext = f"0 {dims[0] - 1} 0 {dims[1] - 1} 0 {dims[2] - 1}"
    o = [float(x) for x in grid.AABB_min]
    xml = (
        '<?xml version="1.0" encoding="utf-8"?>\n'
        '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" '
        'header_type="UInt64" compressor="vtkZLibDataCompressor">\n'
        f'  <ImageData WholeExtent="{ext}" Origin="{o[0]!r} {o[1]!r} {o[2]!r}" '
        f'Spacing="{h!r} {h!r} {h!r}" Direction="1 0 0 0 1 0 0 0 1">\n'
        f'    <Piece Extent="{ext}">\n'
        '      <PointData>\n'
        f'        <DataArray type="{_TYPES[v.dtype]}" Name="{value_label}" format="appended" '
        'offset="0"/>\n'
        '      </PointData>\n'
        '      <CellData>\n'
        '      </CellData>\n'
        '    </Piece>\n'
        '  </ImageData>\n'
        '  <AppendedData encoding="raw">\n_'
    )
    with open(filename, "wb") as f:
        f.write(xml.encode())
        f.write(header)
        for c in comp:
            f.write(c)
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")
    return filename


def read_vti(filename):
    """Reads back a file written by exportSdfToVTI: (values, dims, origin, spacing, name)."""
    import re
    data = open(filename, "rb").read()
    head, _, rest = data.partition(b"<AppendedData encoding=\"raw\">\n_")
    txt = head.decode()
    ext = [int(x) for x in re.search(r'WholeExtent="([^"]+)"', txt).group(1).split()]
    dims = (ext[1] + 1, ext[3] + 1, ext[5] + 1)
    origin = [float(x) for x in re.search(r'Origin="([^"]+)"', txt).group(1).split()]
    spacing = [float(x) for x in re.search(r'Spacing="([^"]+)"', txt).group(1).split()]
    typ = re.search(r'DataArray type="(\w+)"', txt).group(1)
    name = re.search(r'Name="([^"]+)"', txt).group(1)
    nb = int(np.frombuffer(rest[:8], "<u8")[0])
    hdr = np.frombuffer(rest[:8 * (3 + nb)], "<u8")
    pos = 8 * (3 + nb)
    out = b""
    for k in range(nb):
        out += zlib.decompress(rest[pos:pos + int(hdr[3 + k])])
        pos += int(hdr[3 + k])
    dt = np.float64 if typ == "Float64" else np.float32
    return np.frombuffer(out, dt), dims, origin, spacing, name
