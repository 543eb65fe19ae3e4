"""Minimal FITS image reader for beams and maps (host-side ingestion).

Reads the subset of FITS that the reference's beam and map readers consume
(src/beam.cc:295-380, src/fitsData.cc:56-315): 2880-byte blocks of 80-column
header cards, IMAGE HDUs with BITPIX 8/16/32/64/-32/-64, BZERO/BSCALE, and
named extensions (EXTNAME, matched case-insensitively like the reference's
"image"/"IMAGE" and "mask"/"MASK" lookups).  Pixels come back in FITS storage
order (NAXIS1 fastest), which is the order the reference fills its arrays in.
"""
import numpy as np

_BLOCK = 2880
_DTYPES = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


def _parse_value(raw):
    v = raw.split("/", 1)[0].strip() if not raw.strip().startswith("'") \
        else raw.strip()
    if v.startswith("'"):
        end = v.find("'", 1)
        while end != -1 and end + 1 < len(v) and v[end + 1] == "'":
            end = v.find("'", end + 2)
        return v[1:end].replace("''", "'").rstrip()
    if v in ("T", "F"):
        return v == "T"
    try:
        return int(v)
    except ValueError:
        pass
    try:
        return float(v.replace("D", "E"))
    except ValueError:
        return v


class HDU:
    def __init__(self, header, data):
        self.header = header
        self.data = data

    @property
    def name(self):
        return str(self.header.get("EXTNAME", "")).strip()


def read_hdus(path):
    """Return all HDUs of a FITS file as a list of HDU(header, data)."""
    with open(path, "rb") as fh:
        buf = fh.read()
    hdus = []
    pos = 0
    while pos + _BLOCK <= len(buf):
        header = {}
        done = False
        while not done:
            block = buf[pos:pos + _BLOCK]
            pos += _BLOCK
            if len(block) < _BLOCK:
                raise ValueError("truncated FITS header in %s" % path)
            for i in range(0, _BLOCK, 80):
                card = block[i:i + 80].decode("ascii", "replace")
                key = card[:8].strip()
                if key == "END":
                    done = True
                    break
                if card[8:10] == "= ":
                    header[key] = _parse_value(card[10:])
        bitpix = int(header.get("BITPIX", 8))
        naxis = int(header.get("NAXIS", 0))
        dims = [int(header.get("NAXIS%d" % (k + 1), 0)) for k in range(naxis)]
        gcount = int(header.get("GCOUNT", 1))
        pcount = int(header.get("PCOUNT", 0))
        nelem = int(np.prod(dims)) if naxis > 0 else 0
        nbytes = abs(bitpix) // 8 * gcount * (pcount + nelem) if naxis else 0
        data = None
        xt = str(header.get("XTENSION", "IMAGE")).strip().upper()
        if nelem and xt == "IMAGE" and bitpix in _DTYPES:
            raw = np.frombuffer(buf, dtype=_DTYPES[bitpix], count=nelem,
                                offset=pos)
            arr = raw.astype(np.float64 if bitpix < 0 else np.int64)
            bscale = header.get("BSCALE", 1)
            bzero = header.get("BZERO", 0)
            if bscale != 1 or bzero != 0:
                arr = arr * bscale + bzero
            data = arr.reshape(dims[::-1])
        pos += (nbytes + _BLOCK - 1) // _BLOCK * _BLOCK
        hdus.append(HDU(header, data))
    return hdus


def _split_ext(path):
    """cfitsio extended syntax 'file.fits[k]' selects HDU k (0-based)."""
    if path.endswith("]") and "[" in path:
        base, ext = path[:-1].rsplit("[", 1)
        return base, int(ext)
    return path, None


def read_beam(path):
    """Pixels and PIXSIZE (or PIXSCALE) of a beam image, as beam::readFile.

    Returns (pixels flattened in FITS order, pixsize in arcsec)."""
    base, ext = _split_ext(path)
    hdus = read_hdus(base)
    h = hdus[ext if ext is not None else 0]
    if h.data is None or h.data.ndim != 2:
        raise ValueError("Input BEAM is not 2D")
    pixsize = h.header.get("PIXSIZE", h.header.get("PIXSCALE"))
    if pixsize is None:
        raise ValueError("Unable to find pixel scale information in " + path)
    return np.asarray(h.data, dtype=np.float64).ravel(), float(pixsize)


def read_map(path, ignore_mask=False):
    """Unmasked pixels of a map, as fitsData::readData (no mean subtraction).

    The image comes from the extension named image/IMAGE, else the primary
    HDU; pixels whose MASK value is non-zero are dropped."""
    base, ext = _split_ext(path)
    hdus = read_hdus(base)
    if ext is not None:
        img = hdus[ext]
    else:
        named = [h for h in hdus if h.name.lower() == "image"]
        img = named[0] if named else hdus[0]
    if img.data is None or img.data.ndim != 2:
        raise ValueError("Image is not 2D")
    data = np.asarray(img.data, dtype=np.float64).ravel()
    if not ignore_mask:
        masks = [h for h in hdus if h.name.lower() == "mask"]
        if masks:
            m = masks[0].data
            if m is None or m.shape != img.data.shape:
                raise ValueError("Mask does not match image extent")
            keep = np.asarray(m).ravel() == 0
            if not keep.any():
                raise ValueError("All data masked")
            data = data[keep]
    return data


def _image_and_mask(path):
    base, ext = _split_ext(path)
    hdus = read_hdus(base)
    if ext is not None:
        img = hdus[ext]
    else:
        named = [h for h in hdus if h.name.lower() == "image"]
        img = named[0] if named else hdus[0]
    if img.data is None or img.data.ndim != 2:
        raise ValueError("Image is not 2D")
    masks = [h for h in hdus if h.name.lower() == "mask"]
    mask = None
    if masks:
        m = masks[0].data
        if m is None or m.shape != img.data.shape:
            raise ValueError("Mask does not match image extent")
        mask = np.asarray(m).ravel().astype(np.int64).astype(np.uint32)
    return np.asarray(img.data, dtype=np.float64).ravel(), mask


def read_map_pair(path1, path2, ignore_mask=False):
    """Pixel pairs of two maps, as fitsDataDouble::readData
    (src/fitsDataDouble.cc:305-406, no mean subtraction): with both masks
    a pixel is kept when (mask1 & mask2) == 0, with one mask when it is 0."""
    d1, m1 = _image_and_mask(path1)
    d2, m2 = _image_and_mask(path2)
    if d1.size != d2.size:
        raise ValueError("Data from file1 and file2 not the same size")
    if ignore_mask:
        m1 = m2 = None
    if m1 is None and m2 is None:
        return d1, d2
    if m1 is not None and m2 is not None:
        keep = (m1 & m2) == 0
    else:
        keep = (m1 if m1 is not None else m2) == 0
    if not keep.any():
        raise ValueError("All data masked")
    return d1[keep], d2[keep]


def _card(key, value, comment=""):
    if isinstance(value, bool):
        v = "%20s" % ("T" if value else "F")
    elif isinstance(value, int):
        v = "%20d" % value
    elif isinstance(value, float):
        v = "%20.17G" % value
    else:
        v = "'%-8s'" % str(value).replace("'", "''")
    c = "%-8s= %s" % (key, v)
    if comment:
        c += " / " + comment
    return c[:80].ljust(80)


def write_image(path, image, extra=None, extname=None, mask=None):
    """Write a 2-D float64 image (and optionally a MASK extension) to FITS.

    extra: dict of additional header keys for the image HDU (e.g. PIXSIZE).
    With extname the image goes into an IMAGE extension of that name after
    an empty primary HDU."""
    img = np.asarray(image, dtype=">f8")
    assert img.ndim == 2

    def hdu(cards, data):
        head = "".join(cards + ["END".ljust(80)])
        head = head.ljust((len(head) + _BLOCK - 1) // _BLOCK * _BLOCK)
        raw = b"" if data is None else data.tobytes()
        raw += b"\0" * ((-len(raw)) % _BLOCK)
        return head.encode("ascii") + raw

    def img_cards(first, a, bitpix, name=None, keys=None):
        c = [_card("SIMPLE", True)] if first else [_card("XTENSION", "IMAGE")]
        c += [_card("BITPIX", bitpix), _card("NAXIS", 2),
              _card("NAXIS1", a.shape[1]), _card("NAXIS2", a.shape[0])]
        if not first:
            c += [_card("PCOUNT", 0), _card("GCOUNT", 1)]
        if name:
            c.append(_card("EXTNAME", name))
        for k, v in (keys or {}).items():
            c.append(_card(k, v))
        return c

    out = b""
    if extname is None and mask is None:
        out += hdu(img_cards(True, img, -64, keys=extra), img)
    else:
        out += hdu([_card("SIMPLE", True), _card("BITPIX", 8),
                    _card("NAXIS", 0), _card("EXTEND", True)], None)
        out += hdu(img_cards(False, img, -64, extname or "image", extra), img)
        if mask is not None:
            m = np.asarray(mask, dtype=">i2")
            out += hdu(img_cards(False, m, 16, "mask"), m)
    with open(path, "wb") as fh:
        fh.write(out)
