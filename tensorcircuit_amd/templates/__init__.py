from . import blocks, measurements
