"""In-memory stand-in for the subset of the h5py file interface used by the trajectory writers
and System.to_hdf5/from_hdf5 (groups, resizable datasets, attributes). h5py is not installed
here; the writers take any object with this interface."""

import numpy as np


class Dataset(object):
    def __init__(self, shape=None, maxshape=None, dtype=None, data=None):
        self.data = np.array(data) if data is not None else np.zeros(shape, dtype=dtype)
        self.maxshape = maxshape

    @property
    def shape(self):
        return self.data.shape

    def resize(self, n, axis=0):
        assert axis == 0 and self.maxshape is not None and self.maxshape[0] is None
        new = np.zeros((n,) + self.data.shape[1:], dtype=self.data.dtype)
        k = min(n, self.data.shape[0])
        new[:k] = self.data[:k]
        self.data = new

    def __getitem__(self, idx):
        return self.data[idx]

    def __setitem__(self, idx, value):
        self.data[idx] = value


class Group(dict):
    def __init__(self):
        dict.__init__(self)
        self.attrs = {}

    def create_group(self, name):
        assert name not in self
        self[name] = Group()
        return self[name]

    def create_dataset(self, name, shape=None, maxshape=None, dtype=None, data=None):
        assert name not in self
        self[name] = Dataset(shape, maxshape, dtype, data)
        return self[name]
