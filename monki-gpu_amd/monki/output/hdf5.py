"""Read-only access to NetCDF-4 (HDF5) products such as the reference's archived outputs.

netCDF4/h5py are not available here, so this reads the subset of the HDF5 file format
(HDF5 File Format Specification, version 3.0) that the NetCDF-4 library writes for a flat
product: superblock versions 2-3, version 1 and 2 object headers with continuation blocks,
compact and dense (fractal heap + version 2 B-tree) link storage, dataspace, datatype
(fixed-point, floating-point, string), contiguous and compact data layouts, attributes, and
the global heap behind variable-length attributes. NetCDF-4 records each variable's
dimensions as a DIMENSION_LIST attribute of object references to dimension-scale datasets,
whose link names are the dimension names.

``read_netcdf4(path)`` returns ``{name: (array, dimension names, attributes)}`` for every
variable; dimension-scale datasets that are not variables are left out.
"""

from pathlib import Path
from typing import Dict, List, Tuple

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEFINED = 0xFFFFFFFFFFFFFFFF


class HDF5Error(ValueError):
    pass


class _File:
    def __init__(self, data: bytes):
        self.data = data

        if data[:8] != SIGNATURE:
            raise HDF5Error("not an HDF5 file")

        version = data[8]

        if version not in (2, 3):
            raise HDF5Error(f"superblock version {version} is not supported")

        self.offset_size, self.length_size = data[9], data[10]
        position = 12
        self.base = self.address(position)
        position += 2 * self.offset_size  # base address, superblock extension
        position += self.offset_size  # end of file
        self.root = self.address(position)

    # -- primitives --------------------------------------------------------------------

    def unsigned(self, position: int, size: int) -> int:
        return int.from_bytes(self.data[position : position + size], "little")

    def address(self, position: int) -> int:
        return self.unsigned(position, self.offset_size)

    def length(self, position: int) -> int:
        return self.unsigned(position, self.length_size)

    # -- object headers ----------------------------------------------------------------

    def messages(self, address: int) -> List[Tuple[int, bytes]]:
        """All (type, body) messages of the object header at ``address``."""

        data = self.data

        if data[address : address + 4] == b"OHDR":
            return self._messages_v2(address)

        if data[address] == 1:
            return self._messages_v1(address)

        raise HDF5Error(f"unknown object header at {address}")

    def _messages_v2(self, address: int):
        data = self.data
        flags = data[address + 5]
        position = address + 6

        if flags & 0x20:
            position += 16

        if flags & 0x10:
            position += 4

        size_bytes = 1 << (flags & 0x03)
        chunk = self.unsigned(position, size_bytes)
        position += size_bytes
        blocks = [(position, position + chunk)]
        out = []

        while blocks:
            start, end = blocks.pop(0)
            position = start

            while position + 4 <= end:
                kind = data[position]
                size = self.unsigned(position + 1, 2)
                position += 4

                if flags & 0x04:  # attribute creation order tracked: 2-byte order per message
                    position += 2

                body = data[position : position + size]
                position += size

                if kind == 0x10:
                    target = int.from_bytes(body[: self.offset_size], "little")
                    length = int.from_bytes(body[self.offset_size : self.offset_size + self.length_size], "little")

                    if data[target : target + 4] != b"OCHK":
                        raise HDF5Error("bad continuation block")

                    blocks.append((target + 4, target + length - 4))
                elif kind != 0:
                    out.append((kind, body))

        return out

    def _messages_v1(self, address: int):
        data = self.data
        count = self.unsigned(address + 2, 2)
        size = self.unsigned(address + 8, 4)
        blocks = [(address + 16, address + 16 + size)]
        out = []

        while blocks and len(out) < count + 64:
            start, end = blocks.pop(0)
            position = start

            while position + 8 <= end:
                kind = self.unsigned(position, 2)
                length = self.unsigned(position + 2, 2)
                body = data[position + 8 : position + 8 + length]
                position += 8 + length

                if kind == 0x10:
                    target = int.from_bytes(body[: self.offset_size], "little")
                    extent = int.from_bytes(body[self.offset_size : self.offset_size + self.length_size], "little")
                    blocks.append((target, target + extent))
                elif kind != 0:
                    out.append((kind, body))

        return out

    # -- groups ------------------------------------------------------------------------

    def links(self, address: int) -> Dict[str, int]:
        """Hard links of the group at ``address``: name -> object header address."""

        out = {}

        for kind, body in self.messages(address):
            if kind == 0x06:
                name, target = self._link(body)

                if target is not None:
                    out[name] = target
            elif kind == 0x02:
                out.update(self._dense_links(body))
            elif kind == 0x11:
                raise HDF5Error("version 1 (symbol table) groups are not supported")

        return out

    def _link(self, body: bytes):
        flags = body[1]
        position = 2
        link_type = 0

        if flags & 0x08:
            link_type = body[position]
            position += 1

        if flags & 0x04:
            position += 8

        if flags & 0x10:
            position += 1

        size_bytes = 1 << (flags & 0x03)
        name_length = int.from_bytes(body[position : position + size_bytes], "little")
        position += size_bytes
        name = body[position : position + name_length].decode()
        position += name_length

        if link_type != 0:
            return name, None

        return name, int.from_bytes(body[position : position + self.offset_size], "little")

    def _dense_links(self, body: bytes) -> Dict[str, int]:
        flags = body[1]
        position = 2 + (8 if flags & 0x01 else 0)
        heap = int.from_bytes(body[position : position + self.offset_size], "little")
        tree = int.from_bytes(body[position + self.offset_size : position + 2 * self.offset_size], "little")

        if heap == UNDEFINED:
            return {}

        heap_info = self._fractal_heap(heap)
        out = {}

        for heap_id in self._btree_records(tree):
            name, target = self._link(self._heap_object(heap_info, heap_id))

            if target is not None:
                out[name] = target

        return out

    def _fractal_heap(self, address: int) -> dict:
        data = self.data

        if data[address : address + 4] != b"FRHP":
            raise HDF5Error("bad fractal heap header")

        position = address + 5
        id_length = self.unsigned(position, 2)
        filter_length = self.unsigned(position + 2, 2)
        flags = data[position + 4]
        max_managed = self.unsigned(position + 5, 4)
        position += 9
        position += self.length_size  # next huge object id
        position += self.offset_size  # huge object B-tree
        position += self.length_size  # free space
        position += self.offset_size  # free-space manager
        position += 4 * self.length_size  # managed space, allocated, iterator offset, objects
        position += 4 * self.length_size  # huge and tiny object sizes and counts
        table_width = self.unsigned(position, 2)
        starting_block = self.length(position + 2)
        max_direct = self.length(position + 2 + self.length_size)
        max_heap_bits = self.unsigned(position + 2 + 2 * self.length_size, 2)
        position += 4 + 2 * self.length_size + 2
        root = self.address(position)
        rows = self.unsigned(position + self.offset_size, 2)

        if filter_length:
            raise HDF5Error("filtered fractal heaps are not supported")

        return {
            "address": address,
            "id_length": id_length,
            "checksummed": bool(flags & 0x02),
            "max_managed": max_managed,
            "table_width": table_width,
            "starting_block": starting_block,
            "max_direct": max_direct,
            "offset_bytes": (max_heap_bits + 7) // 8,
            "root": root,
            "rows": rows,
        }

    def _heap_object(self, heap: dict, heap_id: bytes) -> bytes:
        if (heap_id[0] >> 4) & 0x03 != 0:
            raise HDF5Error("only managed fractal heap objects are supported")

        offset_bytes = heap["offset_bytes"]
        offset = int.from_bytes(heap_id[1 : 1 + offset_bytes], "little")
        length_bytes = (min(heap["max_direct"], heap["max_managed"]).bit_length() + 7) // 8
        length = int.from_bytes(heap_id[1 + offset_bytes : 1 + offset_bytes + length_bytes], "little")
        block, block_offset = self._direct_block(heap, offset)
        start = block + (offset - block_offset)
        return self.data[start : start + length]

    def _direct_block(self, heap: dict, offset: int) -> Tuple[int, int]:
        """Address and heap offset of the direct block holding heap offset ``offset``."""

        if heap["rows"] == 0:
            return heap["root"], 0

        # Root indirect block: rows of table_width direct blocks, doubling from row 2.
        address = heap["root"]

        if self.data[address : address + 4] != b"FHIB":
            raise HDF5Error("bad fractal heap indirect block")

        position = address + 5 + self.offset_size + heap["offset_bytes"]
        width, size, start = heap["table_width"], heap["starting_block"], 0

        for row in range(heap["rows"]):
            block_size = size if row < 2 else size << (row - 1)

            if block_size > heap["max_direct"]:
                raise HDF5Error("nested fractal heap indirect blocks are not supported")

            for _ in range(width):
                child = self.address(position)
                position += self.offset_size

                if start <= offset < start + block_size and child != UNDEFINED:
                    return child, start

                start += block_size

        raise HDF5Error("fractal heap offset outside the root indirect block")

    def _btree_records(self, address: int) -> List[bytes]:
        data = self.data

        if data[address : address + 4] != b"BTHD":
            raise HDF5Error("bad version 2 B-tree header")

        kind = data[address + 5]
        record_size = self.unsigned(address + 10, 2)
        depth = self.unsigned(address + 12, 2)
        root = self.address(address + 16)
        count = self.unsigned(address + 16 + self.offset_size, 2)

        if kind != 5:
            raise HDF5Error(f"B-tree type {kind} is not a link-name index")

        if depth != 0:
            raise HDF5Error("multi-level link-name B-trees are not supported")

        if data[root : root + 4] != b"BTLF":
            raise HDF5Error("bad B-tree leaf")

        position = root + 6
        records = []

        for _ in range(count):
            record = data[position : position + record_size]
            records.append(record[4:])  # name hash, then heap ID
            position += record_size

        return records

    # -- datasets ----------------------------------------------------------------------

    def dataset(self, address: int):
        shape, dtype, layout, attributes = None, None, None, {}

        for kind, body in self.messages(address):
            if kind == 0x01:
                shape = self._dataspace(body)
            elif kind == 0x03:
                dtype = self._datatype(body)
            elif kind == 0x08:
                layout = body
            elif kind == 0x0C:
                name, value = self._attribute(body)
                attributes[name] = value

        if shape is None or dtype is None or layout is None:
            return None, attributes

        return self._read(layout, shape, dtype), attributes

    def _dataspace(self, body: bytes) -> Tuple[int, ...]:
        version, rank, flags = body[0], body[1], body[2]

        if version == 1:
            position = 8
        elif version == 2:
            position = 4

            if body[3] == 2:  # null dataspace
                return (0,)
        else:
            raise HDF5Error(f"dataspace version {version} is not supported")

        return tuple(
            int.from_bytes(body[position + i * self.length_size : position + (i + 1) * self.length_size], "little")
            for i in range(rank)
        )

    def _datatype(self, body: bytes):
        kind = body[0] & 0x0F
        bits = body[1]
        size = int.from_bytes(body[4:8], "little")
        order = ">" if bits & 0x01 else "<"

        if kind == 0:
            return np.dtype(f"{order}{'i' if bits & 0x08 else 'u'}{size}")

        if kind == 1:
            return np.dtype(f"{order}f{size}")

        if kind == 3:
            return np.dtype(f"S{size}")

        if kind == 7:
            return ("reference", size)

        if kind == 9:
            return ("vlen", self._datatype(body[8:]), bits & 0x0F)

        return ("unsupported", kind)  # compound, enum, ...: skipped in attributes

    def _read(self, layout: bytes, shape, dtype):
        version = layout[0]

        if version not in (3, 4):
            raise HDF5Error(f"data layout version {version} is not supported")

        layout_class = layout[1]
        count = int(np.prod(shape)) if shape else 1

        if not isinstance(dtype, np.dtype):
            raise HDF5Error("only numeric and string datasets are read")

        if layout_class == 0:
            size = int.from_bytes(layout[2:4], "little")
            raw = layout[4 : 4 + size]
        elif layout_class == 1:
            address = int.from_bytes(layout[2 : 2 + self.offset_size], "little")

            if address == UNDEFINED:
                raw = b"\x00" * (count * dtype.itemsize)
            else:
                raw = self.data[address : address + count * dtype.itemsize]
        else:
            raise HDF5Error("chunked datasets are not supported")

        return np.frombuffer(raw, dtype=dtype, count=count).reshape(shape).copy()

    def _attribute(self, body: bytes):
        version = body[0]
        name_size = int.from_bytes(body[2:4], "little")
        type_size = int.from_bytes(body[4:6], "little")
        space_size = int.from_bytes(body[6:8], "little")

        if version == 1:
            position = 8
            pad = lambda n: (n + 7) & ~7  # noqa: E731
        elif version in (2, 3):
            position = 8 + (1 if version == 3 else 0)
            pad = lambda n: n  # noqa: E731
        else:
            raise HDF5Error(f"attribute version {version} is not supported")

        name = body[position : position + name_size].split(b"\x00", 1)[0].decode()
        position += pad(name_size)
        dtype = self._datatype(body[position : position + type_size])
        position += pad(type_size)
        shape = self._dataspace(body[position : position + space_size])
        position += pad(space_size)
        raw = body[position:]
        count = int(np.prod(shape)) if shape else 1

        if isinstance(dtype, np.dtype):
            value = np.frombuffer(raw, dtype=dtype, count=count).reshape(shape)

            if dtype.kind == "S":
                text = [v.split(b"\x00", 1)[0].decode() for v in value.ravel()]
                return name, text[0] if len(text) == 1 else text

            return name, value.copy()

        if dtype[0] == "vlen":
            element = 4 + self.offset_size + 4
            items = [self._vlen(raw[i * element : (i + 1) * element], dtype[1]) for i in range(count)]

            if isinstance(dtype[1], np.dtype) and dtype[2] == 1:  # variable-length string
                return name, items[0] if count == 1 else items

            return name, items

        if dtype[0] == "reference":
            return name, [int.from_bytes(raw[i * self.offset_size : (i + 1) * self.offset_size], "little")
                          for i in range(count)]  # fmt: skip

        return name, None

    def _vlen(self, descriptor: bytes, base):
        length = int.from_bytes(descriptor[:4], "little")
        collection = int.from_bytes(descriptor[4 : 4 + self.offset_size], "little")
        index = int.from_bytes(descriptor[4 + self.offset_size : 8 + self.offset_size], "little")
        raw = self._global_object(collection, index)

        if isinstance(base, tuple) and base[0] == "reference":
            return [int.from_bytes(raw[i * self.offset_size : (i + 1) * self.offset_size], "little")
                    for i in range(length)]  # fmt: skip

        if isinstance(base, np.dtype):
            if base.kind == "S" or base.itemsize == 1:
                return raw[:length].decode(errors="replace")

            return np.frombuffer(raw, dtype=base, count=length).copy()

        return None

    def _global_object(self, collection: int, index: int) -> bytes:
        data = self.data

        if data[collection : collection + 4] != b"GCOL":
            raise HDF5Error("bad global heap collection")

        size = self.length(collection + 8)
        position, end = collection + 8 + self.length_size, collection + size

        while position + 8 + self.length_size <= end:
            number = self.unsigned(position, 2)
            object_size = self.length(position + 8)
            start = position + 8 + self.length_size

            if number == index:
                return data[start : start + object_size]

            if number == 0:
                break

            position = start + ((object_size + 7) & ~7)

        raise HDF5Error(f"global heap object {index} not found")


def read_netcdf4(path) -> Dict[str, Tuple[np.ndarray, Tuple[str, ...], dict]]:
    """Variables of a flat NetCDF-4 file: ``{name: (array, dimension names, attributes)}``."""

    hdf = _File(Path(path).read_bytes())
    links = hdf.links(hdf.root)
    names_by_address = {address: name for name, address in links.items()}
    datasets = {name: hdf.dataset(address) for name, address in links.items()}
    out = {}

    for name, (value, attributes) in datasets.items():
        if value is None:
            continue

        scale = attributes.get("CLASS") == "DIMENSION_SCALE"
        label = attributes.get("NAME", "")

        if scale and isinstance(label, str) and label.startswith("This is a netCDF dimension but not a netCDF variable"):
            continue

        if scale:
            dimensions = (name,)
        elif "DIMENSION_LIST" in attributes:
            dimensions = tuple(names_by_address[refs[0]] for refs in attributes["DIMENSION_LIST"])
        else:
            dimensions = ()

        kept = {
            key: (item.item() if isinstance(item, np.ndarray) and item.size == 1 else item)
            for key, item in attributes.items()
            if key not in ("CLASS", "NAME", "DIMENSION_LIST", "REFERENCE_LIST", "_Netcdf4Dimid", "_Netcdf4Coordinates")
        }
        out[name] = (value, dimensions, kept)

    return out
