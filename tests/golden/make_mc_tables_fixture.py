"""Write tests/golden/mc_tables.json from the reference's marching-cubes
tables (meshing/mc_tables.h): the edge masks, the triangle rows, the corner
offsets, the edge endpoints and the edge owners, as plain data.

Run once in a container that has the reference checkout:
    python tests/golden/make_mc_tables_fixture.py /root/reference
The tests read only the JSON it writes.
"""
import json
import re
import sys
from pathlib import Path


def arrays(text, name):
    body = text[text.index(name):]
    body = body[body.index("{") + 1:body.index("};")]
    return body


def main(ref_root: str) -> None:
    src = (Path(ref_root) / "src/meshing/mc_tables.h").read_text()
    edges = [int(x, 16) for x in re.findall(r"0x[0-9a-fA-F]+", arrays(src, "kCubeEdges[256]"))]
    rows = [[int(v) for v in r.split(",")]
            for r in re.findall(r"\{([^{}]*)\}", arrays(src, "kTriangleVertexEdge[256][16]"))]
    tri = [[e for e in r if e >= 0] for r in rows]

    def small(name):
        return [[int(v) for v in r.split(",")]
                for r in re.findall(r"\{([^{}]*)\}", arrays(src, name))]

    out = {
        "source": "meshing/mc_tables.h",
        "cube_edges": edges,
        "triangles": tri,
        "corner_offsets": small("kVtxOffset[8]"),
        "edge_endpoints": small("kEdgeEndpointVertices[12]"),
        "edge_owner_offset_axis": small("kEdgeOwnerCubeOffset[12]"),
    }
    assert len(edges) == 256 and len(tri) == 256
    Path(__file__).with_name("mc_tables.json").write_text(json.dumps(out, separators=(",", ":")))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
