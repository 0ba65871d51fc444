"""Writes tests/golden/cbox_samples.json from the reference's two sample
renders of Asset/cbox.json ("Sample - Path Tracing.png" and "Sample - Primary
Sample Space Metropolis Light Transport.png", 1024x768 RGBA, 8 bits).

The fixture holds, for each image, the mean RGB of all pixels and a 32 x 24
thumbnail of 32 x 32-pixel block means, both as the 8-bit values / 255.  The
images are the Rgba32 output of Film.Save (Base/Film.fs:55-66): the tone
mapped (ACES, cbox.json's film) and clamped linear film, quantised to bytes,
rows stored top first in Film.Pixels order.  Their sample counts and seeds are
not recorded in the reference, so they pin a converged render only weakly.

Run once in a container where the reference is mounted (reads
/root/reference); the PNGs are decoded with Pillow.
"""
import json
import os

import numpy as np
from PIL import Image

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = {"path_tracing": "Sample - Path Tracing.png",
         "pssmlt": "Sample - Primary Sample Space Metropolis Light Transport.png"}


def main():
    out = {"source": "reference sample renders of Asset/cbox.json, 8-bit RGBA, 1024x768",
           "thumbnail_block": 32}
    for key, name in FILES.items():
        img = np.asarray(Image.open(os.path.join(REF, name)).convert("RGBA"), np.float64)
        assert img.shape == (768, 1024, 4)
        rgb = img[..., :3] / 255.0
        thumb = rgb.reshape(24, 32, 32, 32, 3).mean(axis=(1, 3))
        out[key] = {"mean_rgb": [round(float(v), 6) for v in rgb.mean(axis=(0, 1))],
                    "thumbnail": np.round(thumb, 4).tolist()}
    with open(os.path.join(HERE, "cbox_samples.json"), "w") as fh:
        json.dump(out, fh)


if __name__ == "__main__":
    main()
