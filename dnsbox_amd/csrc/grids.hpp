// The non-cubic grids (NX, NY, NZ) built into the library, beside the cubic grids 16^3 .. 512^3.
// Every pass of a box is instantiated for its grid (the transform lengths and tile counts are
// compile-time constants), so the library carries a fixed list rather than every combination of
// the seven axis lengths: two small grids with three different lengths, a 2:1 box at two
// resolutions (the 2 pi x 2 pi x 4 pi box with alpha_z = 0.5), and a 256 x 256 x 128 grid.
#pragma once

#define DNSBOX_GRIDS(X) X(32, 48, 16) X(48, 16, 32) X(32, 32, 64) X(64, 64, 128) X(256, 256, 128)

namespace dnsbox {

// whether (nx, ny, nz) is one of DNSBOX_GRIDS
inline bool grid_listed(int nx, int ny, int nz)
{
#define DNSBOX_GRID_EQ(a, b, c) || (nx == a && ny == b && nz == c)
    return false DNSBOX_GRIDS(DNSBOX_GRID_EQ);
#undef DNSBOX_GRID_EQ
}

// "32x48x16, 48x16x32, ..." for error messages
inline const char* grid_list_text()
{
#define DNSBOX_GRID_TXT(a, b, c) #a "x" #b "x" #c " "
    return DNSBOX_GRIDS(DNSBOX_GRID_TXT);
#undef DNSBOX_GRID_TXT
}

}  // namespace dnsbox
