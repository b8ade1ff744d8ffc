"""PVP file I/O: OpenPV's binary activity and weight format (host side).

Layout (src/utils/BufferUtilsPvp.hpp:13-99): a header of numParams 4-byte
words -- headerSize, numParams, fileType, nx, ny, nf, numRecords, recordSize,
dataSize, dataType, nxProcs, nyProcs, nxExtended, nyExtended, kx0, ky0,
nBatch, nBands, then a double timestamp -- plus, for weight files, nxp, nyp,
nfp, wMin, wMax (floats) and numPatches.

  fileType 4 (dense activity): per frame a double timestamp and nx*ny*nf floats
  fileType 6 (sparse values):  per frame a double timestamp, an int count and
                               count {int index, float value} pairs (restricted
                               indices, ascending)
  fileType 5 (shared weights): per arbor numPatches x ({u16 nx, u16 ny,
                               u32 offset} + nxp*nyp*nfp floats)
                               (src/io/WeightsFileIO.cpp:147-250)

Values are PV-layout (feature fastest, then x, then y).
"""
import numpy as np

HEADER_WORDS = 20           # 18 ints + one double
WEIGHT_EXTRA_WORDS = 6
PVP_NONSPIKING_ACT_FILE_TYPE = 4
PVP_KERNEL_FILE_TYPE = 5
PVP_ACT_SPARSEVALUES_FILE_TYPE = 6
PVP_FLOAT_TYPE = 3

_NAMES = ["headerSize", "numParams", "fileType", "nx", "ny", "nf", "numRecords", "recordSize",
          "dataSize", "dataType", "nxProcs", "nyProcs", "nxExtended", "nyExtended", "kx0", "ky0",
          "nBatch", "nBands"]


def read_header(raw):
    ints = np.frombuffer(raw[:72], dtype="<i4")
    h = {k: int(v) for k, v in zip(_NAMES, ints)}
    h["timestamp"] = float(np.frombuffer(raw[72:80], dtype="<f8")[0])
    if h["fileType"] == PVP_KERNEL_FILE_TYPE or h["numParams"] > HEADER_WORDS:
        extra = raw[80:80 + 4 * WEIGHT_EXTRA_WORDS]
        ints = np.frombuffer(extra, dtype="<i4")
        floats = np.frombuffer(extra, dtype="<f4")
        h.update(nxp=int(ints[0]), nyp=int(ints[1]), nfp=int(ints[2]), wMin=float(floats[3]),
                 wMax=float(floats[4]), numPatches=int(ints[5]))
    return h


def read_activity(path):
    """Dense (type 4) or sparse-values (type 6) activity: (header, timestamps
    [frames], data [frames][ny][nx][nf] float32)."""
    raw = open(path, "rb").read()
    h = read_header(raw)
    nx, ny, nf = h["nx"], h["ny"], h["nf"]
    n = nx * ny * nf
    off = h["headerSize"]
    times, frames = [], []
    if h["fileType"] == PVP_NONSPIKING_ACT_FILE_TYPE:
        if h["dataType"] != PVP_FLOAT_TYPE:
            raise ValueError("%s: dense activity of data type %d is not float" % (path, h["dataType"]))
        step = 8 + 4 * n
        while off + step <= len(raw):
            times.append(float(np.frombuffer(raw[off:off + 8], dtype="<f8")[0]))
            frames.append(np.frombuffer(raw[off + 8:off + step], dtype="<f4").reshape(ny, nx, nf))
            off += step
    elif h["fileType"] == PVP_ACT_SPARSEVALUES_FILE_TYPE:
        while off + 12 <= len(raw):
            times.append(float(np.frombuffer(raw[off:off + 8], dtype="<f8")[0]))
            cnt = int(np.frombuffer(raw[off + 8:off + 12], dtype="<i4")[0])
            off += 12
            pairs = np.frombuffer(raw[off:off + 8 * cnt],
                                  dtype=np.dtype([("index", "<i4"), ("value", "<f4")]))
            f = np.zeros(n, np.float32)
            f[pairs["index"]] = pairs["value"]
            frames.append(f.reshape(ny, nx, nf))
            off += 8 * cnt
    else:
        raise ValueError("%s: file type %d is not an activity file" % (path, h["fileType"]))
    return h, np.array(times), np.stack(frames).astype(np.float32)


def read_weights(path, all_frames=False):
    """Shared weights (type 5): (header, patches [arbors][numPatches] of (nx, ny,
    offset), weights [arbors][numPatches][nyp][nxp][nfp] float32) of the first
    frame; with all_frames, (headers, timestamps, weights [frames][arbors]...)
    of every frame (each frame carries its own header)."""
    raw = open(path, "rb").read()
    off, headers, times, frames, geos = 0, [], [], [], []
    while off + 80 <= len(raw):
        h = read_header(raw[off:])
        if h["fileType"] != PVP_KERNEL_FILE_TYPE:
            raise ValueError("%s: file type %d is not a shared-weight file" % (path, h["fileType"]))
        nxp, nyp, nfp, npatch = h["nxp"], h["nyp"], h["nfp"], h["numPatches"]
        psize = nxp * nyp * nfp
        off += h["headerSize"]
        arbors = max(1, h["nBands"])
        geo = np.zeros((arbors, npatch, 3), np.int64)
        w = np.zeros((arbors, npatch, nyp, nxp, nfp), np.float32)
        for a in range(arbors):
            for p in range(npatch):
                pnx, pny = np.frombuffer(raw[off:off + 4], dtype="<u2")
                poff = int(np.frombuffer(raw[off + 4:off + 8], dtype="<u4")[0])
                geo[a, p] = (pnx, pny, poff)
                w[a, p] = np.frombuffer(raw[off + 8:off + 8 + 4 * psize], dtype="<f4").reshape(nyp, nxp, nfp)
                off += 8 + 4 * psize
        headers.append(h)
        times.append(h["timestamp"])
        frames.append(w)
        geos.append(geo)
        if not all_frames:
            break
    if not all_frames:
        return headers[0], geos[0], frames[0]
    return headers, np.array(times), np.stack(frames)


def write_weights(path, weights, times):
    """Shared weights (type 5), one frame per timestamp: weights
    [frames][numPatches][nyp][nxp][nfp] (one arbor; numPatches = the data
    patches, here taken as nf along the feature axis).  Each frame is a weight
    header (buildSharedWeightHeader, src/utils/BufferUtilsPvp.cpp:12-98) and
    every patch as {u16 nxp, u16 nyp, u32 0} plus its values
    (WeightsFileIO::writeSharedWeights, src/io/WeightsFileIO.cpp:268-298)."""
    weights = np.ascontiguousarray(weights, np.float32)
    nfr, npatch, nyp, nxp, nfp = weights.shape
    with open(path, "wb") as f:
        for t, w in zip(times, weights):
            hdr = np.array([104, HEADER_WORDS + WEIGHT_EXTRA_WORDS, PVP_KERNEL_FILE_TYPE, 1, 1, npatch, 1, 0,
                            4, PVP_FLOAT_TYPE, 1, 1, 1, 1, 0, 0, 1, 1], "<i4")
            f.write(hdr.tobytes())
            f.write(np.array([t], "<f8").tobytes())
            f.write(np.array([nxp, nyp, nfp], "<i4").tobytes())
            f.write(np.array([w.min(), w.max()], "<f4").tobytes())
            f.write(np.array([npatch], "<i4").tobytes())
            head = np.zeros(1, np.dtype([("nx", "<u2"), ("ny", "<u2"), ("offset", "<u4")]))
            head["nx"], head["ny"] = nxp, nyp
            for p in range(npatch):
                f.write(head.tobytes())
                f.write(w[p].tobytes())


def write_sparse_activity(path, frames, times):
    """Sparse activity with values (type 6): per frame a double timestamp, an
    int count and {int index, float value} for every nonzero restricted
    value in ascending index order (buildSparseActivityHeader and
    writeSparseFrame, src/utils/BufferUtilsPvp.tpp:155-161, 273-282)."""
    frames = np.ascontiguousarray(frames, np.float32)
    nfr, ny, nx, nf = frames.shape
    entry = np.dtype([("index", "<i4"), ("value", "<f4")])
    hdr = np.array([80, HEADER_WORDS, PVP_ACT_SPARSEVALUES_FILE_TYPE, nx, ny, nf, 1, 0, entry.itemsize,
                    PVP_FLOAT_TYPE, 1, 1, nx, ny, 0, 0, 1, nfr], "<i4")
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([times[0] if len(times) else 0.0], "<f8").tobytes())
        for t, fr in zip(times, frames):
            flat = fr.ravel()
            idx = np.nonzero(flat)[0]
            e = np.zeros(idx.size, entry)
            e["index"], e["value"] = idx, flat[idx]
            f.write(np.array([t], "<f8").tobytes())
            f.write(np.array([idx.size], "<i4").tobytes())
            f.write(e.tobytes())


def write_activity(path, frames, times):
    """Dense float activity (type 4): frames [n][ny][nx][nf], one timestamp each."""
    frames = np.ascontiguousarray(frames, np.float32)
    nfr, ny, nx, nf = frames.shape
    hdr = np.zeros(18, "<i4")
    vals = [80, HEADER_WORDS, PVP_NONSPIKING_ACT_FILE_TYPE, nx, ny, nf, 1, nx * ny * nf, 4,
            PVP_FLOAT_TYPE, 1, 1, nx, ny, 0, 0, 1, nfr]
    hdr[:] = vals
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([times[0] if len(times) else 0.0], "<f8").tobytes())
        for t, fr in zip(times, frames):
            f.write(np.array([t], "<f8").tobytes())
            f.write(fr.tobytes())
