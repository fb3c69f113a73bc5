"""File formats: MRC2014 images / volumes (``utils/data_io.py`` counterpart)."""

from .mrc import (  # noqa: F401
    load_mrc_image,
    load_mrc_volume,
    load_template_tensor,
    read_mrc,
    read_mrc_header,
    read_mrc_to_numpy,
    read_mrc_to_tensor,
    write_mrc_from_numpy,
    write_mrc_from_tensor,
)
