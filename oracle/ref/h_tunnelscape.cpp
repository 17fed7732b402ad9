/* harness over code/tunnelscape.cpp */
#include "code/tunnelscape.cpp"
#include "ref_common.h"

/* code/tunnelscape.cpp:44-100 over caller-supplied 2048^2 maps / fog LUT */
REF_API void ref_tscape_ray(uint32_t *pRow, const uint8_t *heightMap,
                            const uint32_t *colorMap, const uint32_t *fog256,
                            int curX, int curY, int dX, int dY)
{
    s_pHeightMap = const_cast<uint8_t *>(heightMap);
    s_pColorMap = const_cast<uint32_t *>(colorMap);
    for (int i = 0; i < 256; ++i) s_fogGradientUnp[i] = c2vISSE16(fog256[i]);
    tscape_ray(pRow, curX, curY, dX, dY);
}
