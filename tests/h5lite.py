"""A small reader for the HDF5 files the reference writes (test infrastructure).

Written from the HDF5 file format specification (version 0 superblock, v1 object
headers, symbol-table groups with v1 B-trees, contiguous and chunked dataset
layouts without filters, fixed-size attributes).  It is used to read the
reference's own result files (heart/test/data/*.h5) and the files this
repository's writer produces, so that the writer is checked by a reader that
also reads files written by the real HDF5 library.
"""
import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF


class H5Error(RuntimeError):
    pass


class Dataset:
    def __init__(self, shape, dtype, layout, attrs):
        self.shape = shape
        self.dtype = dtype
        self.layout = layout
        self.attrs = attrs


class File:
    def __init__(self, path):
        with open(path, 'rb') as f:
            self.buf = f.read()
        b = self.buf
        if b[:8] != b'\x89HDF\r\n\x1a\n':
            raise H5Error('not an HDF5 file')
        if b[8] != 0:
            raise H5Error('superblock version %d not supported' % b[8])
        self.so, self.sl = b[13], b[14]
        if self.so != 8 or self.sl != 8:
            raise H5Error('only 8-byte offsets and lengths')
        self.leaf_k, self.internal_k = struct.unpack_from('<HH', b, 16)
        self.base, _, self.eof, _ = struct.unpack_from('<QQQQ', b, 24)
        # root group symbol table entry at 56
        _, root_oh, cache, _, btree, heap = struct.unpack_from('<QQIIQQ', b, 56)
        self.root = self._group(root_oh, cache, btree, heap)

    # --- low level -------------------------------------------------------
    def _u(self, fmt, off):
        return struct.unpack_from('<' + fmt, self.buf, off)

    def _heap(self, addr):
        if self.buf[addr:addr + 4] != b'HEAP':
            raise H5Error('bad local heap at %d' % addr)
        size, free, data = self._u('QQQ', addr + 8)
        return data, size

    def _heap_name(self, heap, off):
        data, size = self._heap(heap)
        end = self.buf.index(b'\0', data + off)
        return self.buf[data + off:end].decode()

    def _messages(self, addr):
        """Messages of a v1 object header, following continuation blocks."""
        b = self.buf
        if b[addr] != 1:
            raise H5Error('object header version %d at %d' % (b[addr], addr))
        nmsg, = self._u('H', addr + 2)
        hsize, = self._u('I', addr + 8)
        blocks = [(addr + 16, hsize)]
        out = []
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            p = start
            while p + 8 <= start + length and len(out) < nmsg:
                mtype, msize, flags = self._u('HHB', p)
                body = p + 8
                if mtype == 0x10:
                    caddr, clen = self._u('QQ', body)
                    blocks.append((caddr, clen))
                out.append((mtype, body, msize))
                p = body + msize
        return out

    def _group(self, oh, cache, btree, heap):
        if cache != 1:
            for mtype, body, size in self._messages(oh):
                if mtype == 0x11:
                    btree, heap = self._u('QQ', body)
                    break
            else:
                raise H5Error('group without a symbol table')
        entries = {}
        self._walk_group_tree(btree, heap, entries)
        return entries

    def _walk_group_tree(self, addr, heap, entries):
        b = self.buf
        if b[addr:addr + 4] != b'TREE':
            raise H5Error('bad B-tree node at %d' % addr)
        ntype, level = b[addr + 4], b[addr + 5]
        used, = self._u('H', addr + 6)
        p = addr + 24 + 8          # skip key 0
        for _ in range(used):
            child, = self._u('Q', p)
            p += 16                 # child then next key
            if level > 0:
                self._walk_group_tree(child, heap, entries)
            else:
                self._snod(child, heap, entries)

    def _snod(self, addr, heap, entries):
        if self.buf[addr:addr + 4] != b'SNOD':
            raise H5Error('bad symbol node at %d' % addr)
        n, = self._u('H', addr + 6)
        p = addr + 8
        for _ in range(n):
            name_off, oh, cache, _ = self._u('QQII', p)
            entries[self._heap_name(heap, name_off)] = oh
            p += 40

    # --- datatypes and dataspaces ----------------------------------------
    def _datatype(self, body):
        b = self.buf
        cls = b[body] & 0x0F
        bits = b[body + 1:body + 4]
        size, = self._u('I', body + 4)
        if cls == 1:
            order = '>' if bits[0] & 1 else '<'
            return np.dtype(order + ('f%d' % size))
        if cls == 0:
            order = '>' if bits[0] & 1 else '<'
            signed = bool(bits[0] & 8)
            return np.dtype(order + ('i' if signed else 'u') + str(size))
        if cls == 3:
            return np.dtype('S%d' % size)
        raise H5Error('datatype class %d not supported' % cls)

    def _dataspace(self, body):
        b = self.buf
        ver, rank, flags = b[body], b[body + 1], b[body + 2]
        p = body + (8 if ver == 1 else 4)
        dims = self._u('Q' * rank, p)
        return tuple(dims)

    def _attribute(self, body):
        b = self.buf
        ver = b[body]
        nsz, tsz, ssz = self._u('HHH', body + 2)
        if ver != 1:
            raise H5Error('attribute message version %d' % ver)

        def pad(x):
            return (x + 7) // 8 * 8
        p = body + 8
        name = b[p:p + nsz].split(b'\0')[0].decode()
        p += pad(nsz)
        dt = self._datatype(p)
        p += pad(tsz)
        shape = self._dataspace(p)
        p += pad(ssz)
        count = int(np.prod(shape)) if shape else 1
        val = np.frombuffer(b, dtype=dt, count=count, offset=p)
        if dt.kind == 'S':
            val = np.array([v.split(b'\0')[0].decode() for v in val])
        return name, val.reshape(shape) if shape else val[0]

    # --- datasets --------------------------------------------------------
    def dataset(self, name):
        oh = self.root[name]
        shape = dtype = layout = None
        attrs = {}
        for mtype, body, size in self._messages(oh):
            if mtype == 0x01:
                shape = self._dataspace(body)
            elif mtype == 0x03:
                dtype = self._datatype(body)
            elif mtype == 0x08:
                layout = body
            elif mtype == 0x0B:
                raise H5Error('filtered datasets not supported')
            elif mtype == 0x0C:
                k, v = self._attribute(body)
                attrs[k] = v
        return Dataset(shape, dtype, layout, attrs)

    def read(self, name):
        d = self.dataset(name)
        b = self.buf
        body = d.layout
        ver = b[body]
        n = int(np.prod(d.shape))
        if ver in (1, 2):
            # version 1/2: dimensionality, class, 5 reserved bytes, the
            # address (not for compact storage), then the dimension sizes
            # (chunk sizes for chunked storage, element size last)
            ndim, cls = b[body + 1], b[body + 2]
            if cls == 1:
                addr, = self._u('Q', body + 8)
                if addr == UNDEF:
                    return np.zeros(d.shape, dtype=d.dtype)
                return np.frombuffer(b, dtype=d.dtype, count=n, offset=addr).reshape(d.shape).copy()
            if cls == 2:
                btree, = self._u('Q', body + 8)
                cdims = self._u('I' * ndim, body + 16)[:-1]
                out = np.zeros(d.shape, dtype=d.dtype)
                if btree != UNDEF:
                    self._chunks(btree, len(cdims), cdims, d, out)
                return out
            raise H5Error('layout class %d not supported' % cls)
        cls = b[body + 1]
        if ver != 3:
            raise H5Error('layout message version %d' % ver)
        if cls == 1:
            addr, size = self._u('QQ', body + 2)
            if addr == UNDEF:
                return np.zeros(d.shape, dtype=d.dtype)
            return np.frombuffer(b, dtype=d.dtype, count=n, offset=addr).reshape(d.shape).copy()
        if cls == 0:
            size, = self._u('H', body + 2)
            return np.frombuffer(b, dtype=d.dtype, count=n, offset=body + 4).reshape(d.shape).copy()
        if cls == 2:
            rank1 = b[body + 2]
            btree, = self._u('Q', body + 3)
            cdims = self._u('I' * rank1, body + 11)[:-1]
            out = np.zeros(d.shape, dtype=d.dtype)
            if btree != UNDEF:
                self._chunks(btree, len(cdims), cdims, d, out)
            return out
        raise H5Error('layout class %d not supported' % cls)

    def _chunks(self, addr, rank, cdims, d, out):
        b = self.buf
        if b[addr:addr + 4] != b'TREE':
            raise H5Error('bad chunk B-tree at %d' % addr)
        level = b[addr + 5]
        used, = self._u('H', addr + 6)
        ksize = 8 + 8 * (rank + 1)
        p = addr + 24
        for e in range(used):
            csize, fmask = self._u('II', p)
            offs = self._u('Q' * rank, p + 8)
            child, = self._u('Q', p + ksize)
            if level > 0:
                self._chunks(child, rank, cdims, d, out)
            else:
                chunk = np.frombuffer(b, dtype=d.dtype, count=int(np.prod(cdims)), offset=child).reshape(cdims)
                sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, cdims, d.shape))
                out[sl] = chunk[tuple(slice(0, s.stop - s.start) for s in sl)]
            p += ksize + 8
