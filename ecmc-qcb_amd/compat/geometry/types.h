// types.h -- drop-in for src/geometry/types.h (neighbour direction, site parity).
#pragma once

enum dir { up, down };

enum site_parity { EV, OD };
