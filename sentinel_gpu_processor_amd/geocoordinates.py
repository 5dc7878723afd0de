"""Drop-in for ``s1proc.geocoordinates.GeoCoordinates``: the ROI_PAC-style
``.rsc`` lat/lon grid (geocoordinates.py:9-185).  Host metadata code."""

import os

import numpy as np

RA = 6371000.0  # spherical radius used for pixel spacings (geocoordinates.py:7)


def _parse_rsc(path):
    rsc = {}
    with open(path, "r") as f:
        for line in f:
            parts = line.split()
            if len(parts) < 2:
                continue
            key, val = parts[0], parts[1]
            try:
                rsc[key] = float(val) if ("." in val or "e" in val.lower()) else int(val)
            except ValueError:
                rsc[key] = val
    return rsc


class GeoCoordinates:
    def __init__(self, rscfile=None, rscparams=None, metadatafile=None):
        rsc = {}
        if rscfile is not None:
            rsc = _parse_rsc(os.path.join(os.getcwd(), rscfile))
            self.nlon = rsc["WIDTH"]
            self.nlat = rsc["FILE_LENGTH"]
            self.lonmin = rsc["X_FIRST"]
            self.latmax = rsc["Y_FIRST"]
            self.dlon = rsc["X_STEP"]
            self.dlat = rsc["Y_STEP"]
        elif rscparams is not None:
            self.nlon = rscparams["nlon"]
            self.nlat = rscparams["nlat"]
            self.lonmin = rscparams["lonmin"]
            self.latmax = rscparams["latmax"]
            self.dlon = rscparams["dlon"]
            self.dlat = rscparams["dlat"]
        elif metadatafile is not None:
            import pandas as pd
            cols = ["Near Start Lat", "Near Start Lon", "Far Start Lat", "Far Start Lon",
                    "Near End Lat", "Near End Lon", "Far End Lat", "Far End Lon"]
            df = pd.read_csv(metadatafile, header=0, index_col=False)[cols]
            lls = [np.median(df[c].to_numpy()) for c in cols]
            lat, lon = lls[0::2], lls[1::2]
            self.latmax, self.lonmin = np.max(lat), np.min(lon)
            latmin, lonmax = np.min(lat), np.max(lon)
            self.dlon, self.dlat = 1.0 / 7200, -1.0 / 7200
            self.nlon = int(np.round((lonmax - self.lonmin) / self.dlon + 1))
            self.nlat = int(np.round((latmin - self.latmax) / self.dlat + 1))
        else:
            raise Exception("rsc/meta file or rsc parameters need to be specified")
        self.lonmax = self.lonmin + self.dlon * (self.nlon - 1)
        self.latmin = self.latmax + self.dlat * (self.nlat - 1)
        self.latspacing = RA * np.pi / 180
        self.lonspacing = RA * np.cos((self.latmin + self.latmax) * np.pi / 180 / 2) * np.pi / 180
        self.dlatspacing = np.abs(self.dlat) * self.latspacing
        self.dlonspacing = self.dlon * self.lonspacing
        self.zoffset = rsc.get("Z_OFFSET", 0)
        self.zscale = rsc.get("Z_SCALE", 1)
        self.projection = rsc.get("PROJECTION", "LATLON")

    def save_as_rsc(self, filename="out.rsc"):
        with open(os.path.join(os.getcwd(), filename), "w") as f:
            f.write("{0: <16}{1:d}\n".format("WIDTH", int(self.nlon)))
            f.write("{0: <16}{1:d}\n".format("FILE_LENGTH", int(self.nlat)))
            f.write("{0: <16}{1:.14f}\n".format("X_FIRST", self.lonmin))
            f.write("{0: <16}{1:.14f}\n".format("Y_FIRST", self.latmax))
            f.write("{0: <16}{1:.14f}\n".format("X_STEP", self.dlon))
            f.write("{0: <16}{1:.14f}\n".format("Y_STEP", self.dlat))
            f.write("{0: <16}{1}\n".format("Z_OFFSET", self.zoffset))
            f.write("{0: <16}{1}\n".format("Z_SCALE", self.zscale))
            f.write("{0: <16}{1}".format("PROJECTION", self.projection))

    def ll2xy(self, lat, lon):
        if isinstance(lat, (tuple, list, np.ndarray)):
            lat = np.asarray(lat)
            lon = np.asarray(lon)
            return (((lat - self.latmax) / self.dlat).astype(int),
                    ((lon - self.lonmin) / self.dlon).astype(int))
        return int((lat - self.latmax) / self.dlat), int((lon - self.lonmin) / self.dlon)

    def xy2ll(self, x, y):
        return self.latmax + x * self.dlat, self.lonmin + y * self.dlon

    def grid(self):
        return np.mgrid[self.latmax:self.latmin:(self.nlat * 1j),
                        self.lonmin:self.lonmax:(self.nlon * 1j)]

    def take_look(self, latlook, lonlook):
        """Multilooked grid whose pixel centres are look-cell centres
        (geocoordinates.py:153-161)."""
        return GeoCoordinates(rscparams={
            "nlat": self.nlat // latlook, "nlon": self.nlon // lonlook,
            "latmax": self.latmax + self.dlat * (latlook - 1.0) / 2,
            "lonmin": self.lonmin + self.dlon * (lonlook - 1.0) / 2,
            "dlat": self.dlat * latlook, "dlon": self.dlon * lonlook})

    def as_grid(self):
        """The target-grid dict the kernels take."""
        return {"nrow": int(self.nlat), "ncol": int(self.nlon), "lat0": float(self.latmax),
                "lon0": float(self.lonmin), "dlat": float(self.dlat), "dlon": float(self.dlon)}

    def __str__(self):
        return ("Min lat: {0:.6f}\nMax lat: {1:.6f}\nMin lon: {2:.6f}\nMax lon: {3:.6f}\n"
                "Pixels in lat: {4:d}\nPixels in lon: {5:d}\nStep in lat: {6:.6f}\n"
                "Step in lon: {7:.6f}\nLat Pixel Spacing: {8:.6f}\nLon Pixel Spacing: {9:.6f}\n"
                ).format(self.latmin, self.latmax, self.lonmin, self.lonmax, int(self.nlat),
                         int(self.nlon), self.dlat, self.dlon, self.dlatspacing, self.dlonspacing)
