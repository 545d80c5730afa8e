"""A minimal reader for the HDF5 files Korg ships and writes (numpy + zlib, no libhdf5).

Korg stores its linelists (``save_linelist``/``read_linelist``, ``src/linelist.jl:785-839``; the
GALAH DR3 list, ``src/linelist.jl:894-922``), Stark profile tables and test data in HDF5.  The
files are written by HDF5.jl with the library's earliest format: superblock version 0, version-1
object headers, groups as symbol tables (version-1 B-tree of symbol-table nodes plus a local heap),
datasets stored contiguous, compact or chunked (version-1 B-tree of chunks), with the deflate,
shuffle and Fletcher-32 filters.  This module reads exactly that subset, following the published
HDF5 file-format specification (version 3.0), and raises ``H5Error`` on anything else (dense link
storage, version-2 B-trees, variable-length data, ...).

Arrays are returned in the file's (C) dimension order: a Julia ``m x n`` matrix written by HDF5.jl
reads back as ``n x m`` (its transpose), as with any C reader.
"""
import zlib

import numpy as np

__all__ = ["H5Error", "H5File"]


class H5Error(Exception):
    pass


_SIG = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF

# object header message types used here
MSG_NIL, MSG_DATASPACE, MSG_LINK_INFO, MSG_DATATYPE, MSG_FILL_OLD, MSG_FILL = 0x0, 0x1, 0x2, 0x3, 0x4, 0x5
MSG_LINK, MSG_LAYOUT, MSG_FILTERS, MSG_ATTRIBUTE, MSG_CONTINUATION, MSG_SYMBOL_TABLE = (
    0x6, 0x8, 0xB, 0xC, 0x10, 0x11)

FILTER_DEFLATE, FILTER_SHUFFLE, FILTER_FLETCHER32 = 1, 2, 3
MAX_DATASET_BYTES = 1 << 34  # 16 GiB: larger shapes mean a corrupt header here


class H5File:
    """Read-only view of an HDF5 file: ``keys()``, ``read(path)``, ``shape(path)``, ``attrs(path)``.

    ``path`` names a dataset by its group path ("wl", "group/sub/data").
    """

    def __init__(self, path):
        with open(path, "rb") as f:
            self._d = f.read()
        d = self._d
        # the superblock may sit at 0, 512, 1024, ... (a user block before it)
        base = 0
        while base < len(d) and d[base:base + 8] != _SIG:
            base = 512 if base == 0 else base * 2
        if base >= len(d):
            raise H5Error(f"{path}: not an HDF5 file")
        v = d[base + 8]
        if v not in (0, 1):
            raise H5Error(f"{path}: superblock version {v} is not supported")
        self.so, self.sl = d[base + 13], d[base + 14]
        if self.so != 8 or self.sl != 8:
            raise H5Error(f"{path}: offset/length sizes {self.so}/{self.sl} are not supported")
        p = base + 24 + (4 if v == 1 else 0)
        self.base = self._u(p, 8) + (base if self._u(p, 8) == 0 else 0)
        p += 4 * 8  # base, free-space, end-of-file, driver-info addresses
        self.root = self._symbol_entry(p)[1]

    # -- primitive reads ----------------------------------------------------------------------
    def _u(self, p, n):
        if p < 0 or p + n > len(self._d):
            raise H5Error(f"read of {n} bytes at {p:#x} is past the end of the file (truncated?)")
        return int.from_bytes(self._d[p:p + n], "little")

    def _bytes(self, p, n):
        if p < 0 or n < 0 or p + n > len(self._d):
            raise H5Error(f"block of {n} bytes at {p:#x} is past the end of the file (truncated?)")
        return self._d[p:p + n]

    def _addr(self, a):
        return a + self.base

    def _symbol_entry(self, p):
        """Symbol-table entry: (link-name heap offset, object header address)."""
        return self._u(p, 8), self._u(p + 8, 8)

    def _cstr(self, p):
        e = self._d.index(b"\0", p)
        return self._d[p:e].decode("utf-8")

    # -- object headers -----------------------------------------------------------------------
    def _messages(self, addr):
        """Messages of a version-1 object header, continuation blocks included."""
        d = self._d
        p = self._addr(addr)
        if d[p] != 1:
            raise H5Error(f"object header version {d[p]} at {addr:#x} is not supported")
        nmsg = self._u(p + 2, 2)
        blocks = [(p + 16, self._u(p + 8, 4))]
        out = []
        while blocks and len(out) < nmsg:
            q, size = blocks.pop(0)
            end = q + size
            while q + 8 <= end and len(out) < nmsg:
                mtype, msize, mflags = self._u(q, 2), self._u(q + 2, 2), d[q + 4]
                body = q + 8
                if mtype == MSG_CONTINUATION:
                    blocks.append((self._addr(self._u(body, 8)), self._u(body + 8, 8)))
                out.append((mtype, mflags, body, msize))
                q = body + msize
        return out

    # -- groups -------------------------------------------------------------------------------
    def _heap_data(self, heap_addr):
        p = self._addr(heap_addr)
        if self._d[p:p + 4] != b"HEAP":
            raise H5Error(f"no local heap at {heap_addr:#x}")
        return self._addr(self._u(p + 24, 8))

    def _walk_group_btree(self, addr, heap, out):
        p = self._addr(addr)
        d = self._d
        if d[p:p + 4] != b"TREE" or d[p + 4] != 0:
            raise H5Error(f"no group B-tree node at {addr:#x}")
        level, used = d[p + 5], self._u(p + 6, 2)
        q = p + 24 + 8  # past the header and the first key
        for _ in range(used):
            child = self._u(q, 8)
            q += 16  # child address, next key
            if level > 0:
                self._walk_group_btree(child, heap, out)
                continue
            s = self._addr(child)
            if d[s:s + 4] != b"SNOD":
                raise H5Error(f"no symbol-table node at {child:#x}")
            for k in range(self._u(s + 6, 2)):
                name_off, obj = self._symbol_entry(s + 8 + 40 * k)
                out[self._cstr(heap + name_off)] = obj

    def _members(self, addr):
        """Name -> object header address of a group's members."""
        out = {}
        for mtype, _, body, size in self._messages(addr):
            if mtype == MSG_SYMBOL_TABLE:
                btree, heap = self._u(body, 8), self._u(body + 8, 8)
                self._walk_group_btree(btree, self._heap_data(heap), out)
            elif mtype == MSG_LINK:
                name, target = self._link(body)
                if target is not None:
                    out[name] = target
            elif mtype == MSG_LINK_INFO:
                fheap = self._u(body + 2 + (8 if self._d[body + 1] & 1 else 0), 8)
                if fheap != UNDEF:
                    raise H5Error("dense link storage (fractal heap) is not supported")
        return out

    def _link(self, p):
        d = self._d
        flags = d[p + 1]
        q = p + 2
        ltype = 0
        if flags & 0x8:
            ltype = d[q]
            q += 1
        if flags & 0x4:
            q += 8
        if flags & 0x10:
            q += 1
        nlen_size = 1 << (flags & 3)
        nlen = self._u(q, nlen_size)
        q += nlen_size
        name = d[q:q + nlen].decode("utf-8")
        q += nlen
        return name, (self._u(q, 8) if ltype == 0 else None)

    def _resolve(self, path):
        addr = self.root
        for part in [x for x in path.split("/") if x]:
            m = self._members(addr)
            if part not in m:
                raise KeyError(path)
            addr = m[part]
        return addr

    def keys(self, group="/"):
        return sorted(self._members(self._resolve(group)))

    # -- datasets -----------------------------------------------------------------------------
    def _dataset_info(self, addr):
        info = {"filters": []}
        for mtype, _, body, size in self._messages(addr):
            if mtype == MSG_DATASPACE:
                info["shape"] = self._dataspace(body)
            elif mtype == MSG_DATATYPE:
                info["dtype"] = self._datatype(body)
            elif mtype == MSG_LAYOUT:
                info["layout"] = self._layout(body)
            elif mtype == MSG_FILTERS:
                info["filters"] = self._filters(body)
        if "layout" not in info:
            raise H5Error("not a dataset")
        return info

    def _dataspace(self, p):
        d = self._d
        version, ndim, flags = d[p], d[p + 1], d[p + 2]
        if version == 1:
            q = p + 8
        elif version == 2:
            if d[p + 3] == 2:  # null dataspace
                return None
            q = p + 4
        else:
            raise H5Error(f"dataspace version {version}")
        return tuple(self._u(q + 8 * i, 8) for i in range(ndim))

    def _datatype(self, p):
        d = self._d
        cls = d[p] & 0x0F
        bits = d[p + 1] | (d[p + 2] << 8) | (d[p + 3] << 16)
        size = self._u(p + 4, 4)
        order = ">" if bits & 1 else "<"
        if cls == 0:  # fixed-point
            return np.dtype(f"{order}{'i' if bits & 0x8 else 'u'}{size}")
        if cls == 1:  # IEEE floating-point
            if size not in (2, 4, 8):
                raise H5Error(f"{size}-byte floating-point type")
            return np.dtype(f"{order}f{size}")
        if cls == 3:  # fixed-length string
            return np.dtype(f"S{size}")
        raise H5Error(f"datatype class {cls} is not supported")

    def _layout(self, p):
        d = self._d
        version = d[p]
        if version == 3:
            cls = d[p + 1]
            q = p + 2
            if cls == 0:
                n = self._u(q, 2)
                return ("compact", q + 2, n)
            if cls == 1:
                return ("contiguous", self._u(q, 8), self._u(q + 8, 8))
            if cls == 2:
                ndim = d[q]
                btree = self._u(q + 1, 8)
                dims = tuple(self._u(q + 9 + 4 * i, 4) for i in range(ndim))
                return ("chunked", btree, dims)
        elif version in (1, 2):
            ndim, cls = d[p + 1], d[p + 2]
            q = p + 8
            if cls == 0:
                dims = tuple(self._u(q + 4 * i, 4) for i in range(ndim))
                q += 4 * ndim
                return ("compact", q + 4, self._u(q, 4))
            addr = self._u(q, 8)
            dims = tuple(self._u(q + 8 + 4 * i, 4) for i in range(ndim))
            if cls == 1:
                return ("contiguous", addr, None)
            return ("chunked", addr, dims)
        raise H5Error(f"data layout version {version} is not supported")

    def _filters(self, p):
        d = self._d
        version, n = d[p], d[p + 1]
        q = p + (8 if version == 1 else 2)
        out = []
        for _ in range(n):
            fid = self._u(q, 2)
            q += 2
            nlen = 0
            if version == 1 or fid >= 256:
                nlen = self._u(q, 2)
                q += 2
            q += 2  # flags
            nval = self._u(q, 2)
            q += 2
            if version == 1:
                q += (nlen + 7) // 8 * 8
            else:
                q += nlen
            vals = [self._u(q + 4 * i, 4) for i in range(nval)]
            q += 4 * nval
            if version == 1 and nval % 2:
                q += 4
            out.append((fid, vals))
        return out

    def _unfilter(self, raw, filters, mask, itemsize):
        for k in range(len(filters) - 1, -1, -1):
            if mask & (1 << k):
                continue
            fid, vals = filters[k]
            if fid == FILTER_DEFLATE:
                raw = zlib.decompress(raw)
            elif fid == FILTER_SHUFFLE:
                a = np.frombuffer(raw, dtype=np.uint8)
                n = a.size // itemsize
                raw = a[:n * itemsize].reshape(itemsize, n).T.tobytes() + a[n * itemsize:].tobytes()
            elif fid == FILTER_FLETCHER32:
                raw = raw[:-4]
            else:
                raise H5Error(f"filter {fid} is not supported")
        return raw

    def _chunks(self, addr, ndim, out):
        """(filter mask, raw bytes, element offsets) of every chunk below B-tree node ``addr``."""
        d = self._d
        p = self._addr(addr)
        if d[p:p + 4] != b"TREE" or d[p + 4] != 1:
            raise H5Error(f"no chunk B-tree node at {addr:#x}")
        level, used = d[p + 5], self._u(p + 6, 2)
        key_size = 8 + 8 * ndim
        q = p + 24
        for _ in range(used):
            size, mask = self._u(q, 4), self._u(q + 4, 4)
            offs = tuple(self._u(q + 8 + 8 * i, 8) for i in range(ndim))
            child = self._u(q + key_size, 8)
            if level > 0:
                self._chunks(child, ndim, out)
            else:
                c = self._addr(child)
                out.append((mask, self._bytes(c, size), offs))
            q += key_size + 8

    def shape(self, path):
        return self._dataset_info(self._resolve(path))["shape"]

    def read(self, path):
        info = self._dataset_info(self._resolve(path))
        shape, dt = info["shape"], info["dtype"]
        if shape is None:
            return np.empty(0, dtype=dt)
        n = 1
        for x in shape:
            n *= x
        if n * dt.itemsize > MAX_DATASET_BYTES:
            raise H5Error(f"{path}: {n} elements of {dt} exceed this reader's limit")
        kind = info["layout"][0]
        if kind == "compact":
            _, q, size = info["layout"]
            return np.frombuffer(self._bytes(q, n * dt.itemsize), dtype=dt).reshape(shape).copy()
        if kind == "contiguous":
            _, addr, _ = info["layout"]
            if addr == UNDEF:  # never written: the fill value (zero)
                return np.zeros(shape, dtype=dt)
            q = self._addr(addr)
            return np.frombuffer(self._bytes(q, n * dt.itemsize), dtype=dt).reshape(shape).copy()
        _, btree, cdims = info["layout"]
        cshape = cdims[:-1]
        if cdims[-1] != dt.itemsize or len(cshape) != len(shape):
            raise H5Error("chunk dimensions do not match the dataset")
        out = np.zeros(shape, dtype=dt)
        if btree == UNDEF:
            return out
        chunks = []
        self._chunks(btree, len(cdims), chunks)
        cn = 1
        for x in cshape:
            cn *= x
        for mask, raw, offs in chunks:
            raw = self._unfilter(raw, info["filters"], mask, dt.itemsize)
            if len(raw) < cn * dt.itemsize:
                raise H5Error(f"{path}: a chunk holds {len(raw)} bytes, {cn * dt.itemsize} expected")
            block = np.frombuffer(raw[:cn * dt.itemsize], dtype=dt).reshape(cshape)
            sel_out, sel_in = [], []
            for o, c, s in zip(offs, cshape, shape):
                e = min(o + c, s)
                sel_out.append(slice(o, e))
                sel_in.append(slice(0, e - o))
            out[tuple(sel_out)] = block[tuple(sel_in)]
        return out

    def attrs(self, path="/"):
        """Scalar or 1-D attributes of numeric or fixed-string type (others are skipped)."""
        addr = self._resolve(path)
        out = {}
        for mtype, _, body, size in self._messages(addr):
            if mtype != MSG_ATTRIBUTE:
                continue
            d = self._d
            version = d[body]
            if version not in (1, 2, 3):
                continue
            nlen, tlen, slen = self._u(body + 2, 2), self._u(body + 4, 2), self._u(body + 6, 2)
            pad = (lambda x: (x + 7) // 8 * 8) if version == 1 else (lambda x: x)
            q = body + (9 if version == 3 else 8)  # version 3 adds the name's character set
            name = d[q:q + nlen].rstrip(b"\0").decode("utf-8")
            q += pad(nlen)
            try:
                dt = self._datatype(q)
            except H5Error:
                continue
            q += pad(tlen)
            shp = self._dataspace(q)
            q += pad(slen)
            n = 1
            for x in shp or ():
                n *= x
            val = np.frombuffer(d[q:q + n * dt.itemsize], dtype=dt)
            out[name] = val[0] if not shp else val.reshape(shp).copy()
        return out
