"""Settings the reference keeps in module globals (cfg.py:1-14) and reads
from options.json, for the paths here that need them: the erosion snapshot
of erode_terrain3 (erosion.py:186-192) and the image export helpers.

SNAP_DIR        cfg.py:5, the folder snapshots are written to
IMG_QUERY_DATA  cfg.py:14, (dists, nbrs) of every pixel's nearest vertices,
                as KDTree.query / KnnGrid.query_pixels_device return them
                (host arrays or device tensors, shaped [height, width, k])
IMG_WIDTH, IMG_HEIGHT, IMG_FORMAT
                options.json img_width / img_height / img_format, with the
                reference's defaults
"""
SNAP_DIR = None
IMG_QUERY_DATA = None
IMG_WIDTH = 2048
IMG_HEIGHT = 1024
IMG_FORMAT = "png"
