"""An independent numpy statement of the Gadget format-2 file as the
reference writes it (IO/write_snapshot.c:63-297, header IO/io.h:23-42):
blocks of two Fortran records (the 8-byte label record: 4 characters and
the data record's size + 8; then the data record), a Fortran record being
its payload between two 4-byte copies of its length."""
import numpy as np

# struct gadget_header with C's layout rules: every member at the next
# multiple of its own size, the struct padded to a multiple of 8
HEADER = np.dtype({
    "names": ["Npart", "Massarr", "Time", "Redshift", "Flag_Sfr",
              "Flag_Feedback", "Nall", "Flag_Cooling", "Num_Files",
              "Boxsize", "Omega0", "Omega_Lambda", "Hubble_Param",
              "Flag_Age", "Flag_Metals", "Nall_High_Word"],
    "formats": [("<u4", 6), ("<f8", 6), "<f8", "<f8", "<i4", "<i4",
                ("<u4", 6), "<i4", "<i4", "<f8", "<f8", "<f8", "<f8",
                "<i4", "<i4", ("<u4", 6)],
    "offsets": [0, 24, 72, 80, 88, 92, 96, 120, 124, 128, 136, 144, 152,
                160, 164, 168],
    "itemsize": 256})


def header(npart, massarr=(0.0,) * 6, **kw):
    h = np.zeros((), dtype=HEADER)
    h["Npart"] = npart
    h["Nall"] = npart
    h["Massarr"] = massarr
    h["Num_Files"] = 1
    for k, v in kw.items():
        h[k] = v
    return h


def _record(payload, endian="<"):
    n = np.array([len(payload)], dtype=endian + "u4").tobytes()
    return n + payload + n


def write(path, head, blocks, endian="<"):
    """blocks: [(label, array)] after HEAD, each array written as is
    (particles' components together)."""
    with open(path, "wb") as f:
        for label, data in [("HEAD", np.asarray(head))] + list(blocks):
            payload = np.ascontiguousarray(data).tobytes()
            size = np.array([len(payload) + 8], dtype=endian + "u4")
            f.write(_record(label.encode() + size.tobytes(), endian))
            f.write(_record(payload, endian))


def read(path):
    """{label: raw bytes} in file order (dict order), and the header."""
    out = {}
    with open(path, "rb") as f:
        buf = f.read()
    pos = 0
    while pos < len(buf):
        r0 = int(np.frombuffer(buf, "<u4", 1, pos)[0])
        assert r0 == 8
        label = buf[pos + 4:pos + 8].decode()
        nxt = int(np.frombuffer(buf, "<u4", 1, pos + 8)[0])
        assert int(np.frombuffer(buf, "<u4", 1, pos + 12)[0]) == 8
        pos += 16
        size = int(np.frombuffer(buf, "<u4", 1, pos)[0])
        assert size + 8 == nxt
        out[label] = buf[pos + 4:pos + 4 + size]
        assert int(np.frombuffer(buf, "<u4", 1, pos + 4 + size)[0]) == size
        pos += size + 8
    head = np.frombuffer(out["HEAD"], dtype=HEADER)[0]
    return out, head
