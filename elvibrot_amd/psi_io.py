"""Wave-packet files in ElVibRot's formats, so that vectors computed on the
GPU (Davidson eigenvectors, propagated packets) can be read by the Fortran
code (``read_WP=t``, ``name_file_readWP``) and the other way round.

Formats (SRC/sub_WP/sub_module_psi_io.f90):

* version 0, formatted (the default, ``FilePsiVersion = 0``,
  SRC/EVR_system_m.f90:104): header line ``nb_psi nb_tot``
  (Write_header_saveFile_psi :1312-1369), then for every vector one line
  ``i_1 .. i_D i_h i_e a`` (``a b`` for a complex vector) per basis function
  with ``|a| >= epsi``, the ``i_k`` being the 1-based mode indices of the
  packed basis function (Rec_ndim_index, SRC/sub_Basis/sub_module_basis.f90
  :3004-3084, SG type 4 branch :3055-3065), closed by ``end wp <j>``
  (Write_Psi_nDBasis :1478-1560).  The reader accepts a one-integer header
  as well and, like lect_psiBasisRepnotall_nD (:264-), matches the index
  tuples to packed rows and skips functions the basis does not hold.
* version 1, formatted: ``&headerFile`` namelist (Version_File, nb_psi,
  nb_tot), the index list (one ``i_1 .. i_D i_h i_e`` line per packed row,
  Write_list_nDindBasis :1370-1477), then per vector ``T`` and the
  ``nb_tot`` coefficients, list-directed (Write_Psi_nDBasis case 1).  The
  reference's reader for this case (Read_psi_nDBasis, :797-812) reads the
  vector without the leading T its writer puts there; the files here follow
  the writer, and the reader here takes T first.
* version 1, unformatted (sequential records with 4-byte length markers):
  ``'X'``, nb_psi, nb_tot, 1, ``'X'``, nb_tot index records, one record
  ``T, coefficients`` per vector.

* propagation restart (SaveWP_restart / ReadWP_restart,
  SRC/sub_propagation/sub_module_propagation.f90:379-495): ``T nb_WP nb_tot``,
  then every packet's complex coefficients list-directed.

Complex values are written as list-directed complex constants ``(re,im)``;
the readers take the forms list-directed input allows (blanks or commas,
``D`` exponents, repeat counts ``r*c``, values over several records, a
complex constant split over two records, namelist items ``NAME= value``),
with each READ starting on a new record as in Fortran.
Several electronic states (``nb_e`` > 1, the multi-surface operators of
elvibrot_amd.surfaces) put the nb_e components one after the other, nb_ba
rows each, with ``i_e`` their 1-based state index, as the reference's loops
over i_e, i_h, i_b do (Write_Psi_nDBasis :1525-1530); there is no extra
index (``i_h = 1``).
"""
import re
import struct

import numpy as np


def packed_indices(basis):
    """(nb_ba, D) int32: 0-based mode indices of every packed row
    (sg4_setup_packed); SG4Batch-like objects without it are not supported."""
    cached = getattr(basis, "_packed_idx", None)
    if cached is None:
        cached = basis.packed()
        basis._packed_idx = cached
    return cached


def _fmt_real(x):
    return f"{x: .17E}"


def _fmt_cplx(c):
    # a complex constant of list-directed I/O: (re,im)
    return f"({c.real:.17E},{c.imag:.17E})"


_LD_TOKEN = re.compile(r"(?:(\d+)\*)?(\([^()]*\)|[^\s,()]+)")


def _ld_values(text):
    """The values of list-directed input: reals, complex constants ``(re,im)``
    and repeat counts ``r*value``, separated by blanks, commas or line ends;
    Fortran ``D`` exponents are accepted."""
    out = []
    for rpt, tok in _LD_TOKEN.findall(text):
        if tok.startswith("("):
            re_, im_ = tok[1:-1].split(",")
            v = complex(float(re_.strip().replace("D", "E").replace("d", "e")),
                        float(im_.strip().replace("D", "E").replace("d", "e")))
        else:
            v = float(tok.replace("D", "E").replace("d", "e"))
        out.extend([v] * (int(rpt) if rpt else 1))
    return out


def _ld_read(lines, pos, n, what="value"):
    """One list-directed READ of n values from the records ``lines[pos:]``:
    it starts on a new record, takes values from successive records until it
    has n, and the next READ starts on the record after the last one used
    (values left on that record are skipped).  Returns (values, next_pos)."""
    got, carry = [], ""
    while len(got) < n:
        if pos >= len(lines):
            raise ValueError(f"end of file inside a list-directed read of {n} {what}s")
        text = carry + " " + lines[pos]
        pos += 1
        # a complex constant may continue on the next record, before or after
        # its comma (list-directed output splits long ones that way)
        m = _LD_OPEN.search(text)
        carry, text = (text[m.start():], text[:m.start()]) if m else ("", text)
        got.extend(_ld_values(text))
    return got[:n], pos


_LD_OPEN = re.compile(r"(?:\d+\*)?\([^()]*$")
_NML_ITEM = re.compile(r"([A-Za-z_]\w*)\s*=\s*([^,/\s]+)")


def _namelist(lines, pos):
    """The items ``NAME = value`` of a namelist group starting at record pos
    (``&group``, items separated by commas or blanks, over one or more
    records, closed by ``/``); names upper-cased.  Returns (dict, next_pos)."""
    text = []
    while True:
        if pos >= len(lines):
            raise ValueError("end of file inside the &headerFile namelist")
        text.append(lines[pos])
        pos += 1
        if "/" in text[-1]:
            break
    body = " ".join(text).split("/", 1)[0]
    body = re.sub(r"^\s*&\w+", "", body)
    return {k.upper(): v for k, v in _NML_ITEM.findall(body)}, pos


def _as_complex(vals, n):
    """n complex numbers from list-directed values: complex constants, or
    (re, im) pairs of reals as written by older files of this module."""
    if len(vals) >= n and all(isinstance(v, complex) for v in vals[:n]):
        return np.array(vals[:n], dtype=np.complex128), n
    r = np.array(vals[:2 * n], dtype=np.float64)
    return r[0::2] + 1j * r[1::2], 2 * n


def _index_list(basis, nb_e):
    """(nb_e nb_ba, D + 2): i_1 .. i_D, i_h = 1, i_e of every row."""
    idx = packed_indices(basis) + 1
    nb = idx.shape[0]
    out = np.ones((nb_e * nb, idx.shape[1] + 2), dtype=np.int64)
    for e in range(nb_e):
        out[e * nb:(e + 1) * nb, :-2] = idx
        out[e * nb:(e + 1) * nb, -1] = e + 1
    return out


def write_psi(path, psi, basis, version=0, epsi=0.0, T=0.0, formatted=True, nb_e=1):
    """Write the vectors ``psi`` (nb_e nb_ba x n, real or complex) as a
    wave-packet file (sub_save_psi, SRC/sub_WP/sub_module_psi_io.f90:65-124)."""
    psi = np.asarray(psi)
    if psi.ndim == 1:
        psi = psi[:, None]
    nb, n = psi.shape
    if nb != nb_e * basis.nb_ba:
        raise ValueError("psi rows do not match the basis and the number of states")
    full = _index_list(basis, nb_e)
    cplx = np.iscomplexobj(psi)
    if not formatted:
        if version != 1:
            raise ValueError("unformatted files are written with version 1 only")
        with open(path, "wb") as f:
            def rec(b):
                f.write(struct.pack("<i", len(b)) + b + struct.pack("<i", len(b)))
            rec(b"X")
            rec(struct.pack("<i", n))
            rec(struct.pack("<i", nb))
            rec(struct.pack("<i", 1))
            rec(b"X")
            fi = np.ascontiguousarray(full.astype("<i4"))
            for r in range(nb):
                rec(fi[r].tobytes())
            for j in range(n):
                v = psi[:, j].astype("<c16" if cplx else "<f8")
                rec(struct.pack("<d", T) + v.tobytes())
        return
    with open(path, "w") as f:
        if version == 0:
            f.write(f" {n} {nb}\n")
            for j in range(n):
                v = psi[:, j]
                keep = np.abs(v) >= epsi
                for r in np.nonzero(keep)[0]:
                    ind = " ".join(str(int(i)) for i in full[r])
                    if cplx:
                        f.write(f" {ind} {_fmt_real(v[r].real)} {_fmt_real(v[r].imag)}\n")
                    else:
                        f.write(f" {ind} {_fmt_real(v[r])}\n")
                f.write(f" end wp {j + 1}\n")
        elif version == 1:
            f.write(f"&HEADERFILE\n VERSION_FILE=1,\n NB_PSI={n},\n NB_TOT={nb},\n /\n")
            for r in range(nb):
                f.write(" " + " ".join(str(int(i)) for i in full[r]) + "\n")
            for j in range(n):
                v = psi[:, j]
                if cplx:
                    f.write(" " + _fmt_real(T) + "\n")
                    for a in range(0, nb, 2):
                        f.write(" " + " ".join(_fmt_cplx(c) for c in v[a:a + 2]) + "\n")
                else:
                    vals = [T] + list(v)
                    for a in range(0, len(vals), 4):
                        f.write(" " + " ".join(_fmt_real(x) for x in vals[a:a + 4]) + "\n")
        else:
            raise ValueError("version must be 0 or 1")


def _rows_of(basis, nb_e=1):
    """(i_1 .. i_D, i_e) -> row of the nb_e-state packed vector"""
    full = _index_list(basis, nb_e)
    return {tuple(int(i) for i in r[:-2]) + (int(r[-1]),): n for n, r in enumerate(full)}


def read_psi(path, basis, cplx=False, formatted=True, nb_e=1):
    """Read a wave-packet file into an (nb_e nb_ba, n) array (complex when
    ``cplx``).  Returns (psi, info) with info = dict(version, nb_psi, nb_tot,
    T (list, version 1), skipped (functions absent from the basis or of a
    state beyond nb_e))."""
    D = basis.D
    nbr = nb_e * basis.nb_ba
    dt = np.complex128 if cplx else np.float64
    info = dict(skipped=0, T=[])
    if not formatted:
        with open(path, "rb") as f:
            data = f.read()
        recs, p = [], 0
        while p < len(data):
            (ln,) = struct.unpack_from("<i", data, p)
            recs.append(data[p + 4:p + 4 + ln])
            p += 8 + ln
        if recs[0] != b"X":
            raise ValueError("not a version-1 unformatted wave-packet file")
        n = struct.unpack("<i", recs[1])[0]
        nb = struct.unpack("<i", recs[2])[0]
        info.update(version=struct.unpack("<i", recs[3])[0], nb_psi=n, nb_tot=nb)
        lst = [np.frombuffer(recs[5 + r], dtype="<i4") for r in range(nb)]
        rows = _rows_of(basis, nb_e)
        target = np.array([rows.get(tuple(int(i) for i in x[:D]) + (int(x[D + 1]),), -1)
                           for x in lst])
        psi = np.zeros((nbr, n), dtype=dt)
        for j in range(n):
            b = recs[5 + nb + j]
            info["T"].append(struct.unpack_from("<d", b, 0)[0])
            v = np.frombuffer(b, dtype="<c16" if cplx else "<f8", offset=8)
            ok = target >= 0
            psi[target[ok], j] = v[ok]
            info["skipped"] += int((~ok).sum())
        return psi, info
    with open(path) as f:
        lines = f.read().splitlines()
    first = lines[0].split()
    rows = _rows_of(basis, nb_e)
    try:
        ints = [int(x) for x in first]
    except ValueError:
        ints = None
    if ints:
        n = ints[0]
        nb = ints[1] if len(ints) > 1 else 0
        info.update(version=0, nb_psi=n, nb_tot=nb)
        psi = np.zeros((nbr, n), dtype=dt)
        j, pos = 0, 1
        while j < n and pos < len(lines):
            line = lines[pos]
            pos += 1
            if "end" in line:
                j += 1
                continue
            tok = line.split()
            if not tok:
                continue
            ind = tuple(int(x) for x in tok[:D]) + (int(tok[D + 1]),)
            vals = [float(x.replace("D", "E").replace("d", "e")) for x in tok[D + 2:]]
            r = rows.get(ind)
            if r is None:
                info["skipped"] += 1
                continue
            psi[r, j] = complex(vals[0], vals[1]) if cplx else vals[0]
        return psi, info
    # version 1: &headerFile namelist, possibly over several lines
    if not lines[0].lstrip().startswith("&"):
        raise ValueError("unrecognised wave-packet file header")
    kv, pos = _namelist(lines, 0)
    n, nb = int(kv["NB_PSI"]), int(kv["NB_TOT"])
    info.update(version=int(kv.get("VERSION_FILE", 1)), nb_psi=n, nb_tot=nb)
    # one READ per packed row: ind(1:D), i_h, i_e
    lst = []
    for r in range(nb):
        v, pos = _ld_read(lines, pos, D + 2, "index")
        lst.append(tuple(int(x.real) for x in v[:D]) + (int(v[D + 1].real),))
    target = np.array([rows.get(t, -1) for t in lst])
    psi = np.zeros((nbr, n), dtype=dt)
    ok = target >= 0
    for j in range(n):
        # one READ per vector: T, vec(1:nb_tot)
        vals, pos2 = _ld_read(lines, pos, nb + 1, "coefficient")
        if cplx and not all(isinstance(x, complex) for x in vals[1:]):
            # (re, im) pairs of reals (files written before the complex constants)
            vals, pos2 = _ld_read(lines, pos, 2 * nb + 1, "coefficient")
        pos = pos2
        info["T"].append(float(vals[0].real))
        if cplx:
            v, _ = _as_complex(vals[1:], nb)
        else:
            v = np.array([x.real for x in vals[1:]])
        psi[target[ok], j] = v[ok]
        info["skipped"] += int((~ok).sum())
    return psi, info


def write_restart(path, T, psis):
    """Propagation restart file (SaveWP_restart,
    SRC/sub_propagation/sub_module_propagation.f90:379-427; written after every
    WPdeltaT by the march loop, sub_module_propa_march.f90:226): the record
    ``T nb_WP nb_tot``, then every packet's complex coefficients (packed
    basis order) list-directed, each packet starting a new record.
    ``psis``: (nb_WP, nb_tot) complex, numpy or torch."""
    if hasattr(psis, "detach"):
        psis = psis.detach().cpu().numpy()
    psis = np.asarray(psis, dtype=np.complex128)
    if psis.ndim == 1:
        psis = psis[None, :]
    nwp, nb = psis.shape
    with open(path, "w") as f:
        f.write(f" {_fmt_real(T)} {nwp} {nb}\n")
        for j in range(nwp):
            v = psis[j]
            for a in range(0, nb, 2):
                f.write(" " + " ".join(_fmt_cplx(c) for c in v[a:a + 2]) + "\n")


def read_restart(path, nb_wp, nb_tot):
    """ReadWP_restart (sub_module_propagation.f90:429-495): returns (T, psis)
    with psis (nb_wp, nb_tot) complex.  A file that does not start with
    ``T nb_WP nb_tot`` (one list-directed read, over one or more records)
    means no restart, as in the reference (a warning there, T0 = 0):
    (0.0, None).  Sizes different from nb_wp, nb_tot raise, as the reference
    stops."""
    with open(path) as f:
        lines = f.read().splitlines()
    # the header is one list-directed read of three values: like any such
    # read it continues onto the next records until it has them
    head, pos = [], 0
    try:
        while len(head) < 3 and pos < len(lines):
            head.extend(_ld_values(lines[pos]))
            pos += 1
        # a complex constant is not a value a read of reals and integers accepts
        if any(isinstance(v, complex) for v in head[:3]):
            raise ValueError
        T, nwp_f, nb_f = float(head[0]), int(head[1]), int(head[2])
        if head[1] != nwp_f or head[2] != nb_f:
            raise ValueError
    except (ValueError, IndexError, AttributeError):
        return 0.0, None
    if nwp_f != nb_wp or nb_f != nb_tot:
        raise ValueError(f"inconsistent WP size values: file {nwp_f} x {nb_f}, "
                         f"expected {nb_wp} x {nb_tot}")
    psis = np.zeros((nb_wp, nb_tot), dtype=np.complex128)
    # every packet is one list-directed read: it starts on a new record and
    # reads nb_tot complex constants across as many records as they take
    for j in range(nb_wp):
        try:
            got, pos = _ld_read(lines, pos, nb_tot, "coefficient")
        except ValueError:
            raise ValueError(f"restart file ends inside packet {j + 1}") from None
        if not all(isinstance(v, complex) for v in got):
            raise ValueError(f"packet {j + 1}: complex constants (re,im) expected")
        psis[j] = got
    return T, psis
