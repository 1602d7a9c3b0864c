"""DC-resistivity survey objects with SimPEG's attribute names.

`TorchDC2D` reads a survey through `survey.source_list`, each source's
`location_a` / `location_b` (a pole has no B) and `receiver_list`, and each
receiver's `locations_m` / `locations_n` (a pole has no N) and `nD`.  A stock
`simpeg.electromagnetics.static.resistivity.survey.Survey` has exactly these
attributes and works unchanged; the classes below supply them on hosts
without SimPEG.  Receivers measure potential ("volt"), with unit current;
an induced-polarisation survey marks receivers "apparent_chargeability"
(stock SimPEG's name), which the DC engines treat as "volt" and the IP
engines divide by the background voltage.
"""
import numpy as np


def _locs(a):
    return np.atleast_2d(np.asarray(a, dtype=float))


_DATA_TYPES = ("volt", "apparent_chargeability")


def _check_type(data_type):
    if data_type not in _DATA_TYPES:
        raise ValueError(f"data_type must be one of {_DATA_TYPES}, got {data_type!r}")
    return data_type


class _Receiver:
    data_type = "volt"

    @property
    def nD(self):
        return int(self.locations_m.shape[0])


class DipoleReceiver(_Receiver):
    """Potential difference u(M) - u(N) for each row of M and N."""

    def __init__(self, locations_m, locations_n, data_type="volt"):
        self.data_type = _check_type(data_type)
        self.locations_m = _locs(locations_m)
        self.locations_n = _locs(locations_n)
        if self.locations_m.shape != self.locations_n.shape:
            raise ValueError("M and N locations differ in shape")

    @property
    def locations(self):
        return [self.locations_m, self.locations_n]


class PoleReceiver(_Receiver):
    """Potential u(M) for each row of M."""

    locations_n = None

    def __init__(self, locations, data_type="volt"):
        self.data_type = _check_type(data_type)
        self.locations_m = _locs(locations)

    @property
    def locations(self):
        return self.locations_m


class DipoleSource:
    """Unit current in at A and out at B."""

    def __init__(self, receiver_list, location_a, location_b, current=1.0):
        self.receiver_list = list(receiver_list)
        self.location_a = np.asarray(location_a, dtype=float).reshape(-1)
        self.location_b = np.asarray(location_b, dtype=float).reshape(-1)
        self.current = float(current)

    @property
    def location(self):
        return [self.location_a, self.location_b]


class PoleSource:
    """Unit current in at A, return electrode at infinity."""

    location_b = None

    def __init__(self, receiver_list, location, current=1.0):
        self.receiver_list = list(receiver_list)
        self.location_a = np.asarray(location, dtype=float).reshape(-1)
        self.current = float(current)

    @property
    def location(self):
        return self.location_a


class Survey:
    def __init__(self, source_list):
        self.source_list = list(source_list)

    @property
    def nD(self):
        return int(sum(rx.nD for s in self.source_list for rx in s.receiver_list))

    @property
    def nSrc(self):
        return len(self.source_list)


class receivers:                     # noqa: N801 - SimPEG's module-style names
    Dipole = DipoleReceiver
    Pole = PoleReceiver


class sources:                       # noqa: N801
    Dipole = DipoleSource
    Pole = PoleSource
