// Exercises the C++ drop-in templates (include/enblend_amd/enblend.hpp) with
// the reference's template arguments for 16-bit RGB images
// (src/numerictraits.h:173-174): reads an interleaved RGB16 image and an
// alpha from argv[1], builds its Laplacian pyramid through
// enblend::laplacianPyramid, writes every level (planar int32) to argv[2],
// then collapses it with enblend::collapsePyramid and checks that level 0 is
// the input << 7 (exact reconstruction).
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "../../include/enblend_amd/enblend.hpp"
#include "mini_image.hpp"

typedef mini::Image<mini::RGB<uint16_t>> ImageType;
typedef mini::Image<uint8_t> AlphaType;
typedef mini::Image<mini::RGB<int32_t>> PyramidType;

int main(int argc, char** argv)
{
    if (argc < 3) { fprintf(stderr, "usage: dropin_test in.bin out.bin\n"); return 2; }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t hdr[3];
    if (fread(hdr, 4, 3, f) != 3) return 2;
    const int w = hdr[0], h = hdr[1], levels = hdr[2];
    ImageType img(w, h);
    AlphaType alpha(w, h);
    std::vector<uint16_t> pix((size_t)w * h * 3);
    std::vector<uint8_t> al((size_t)w * h);
    if (fread(pix.data(), 2, pix.size(), f) != pix.size() || fread(al.data(), 1, al.size(), f) != al.size()) return 2;
    fclose(f);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            img(x, y) = mini::RGB<uint16_t>(pix[3 * i], pix[3 * i + 1], pix[3 * i + 2]);
            alpha(x, y) = al[i];
        }
    try {
        std::vector<PyramidType*>* lp =
            enblend::laplacianPyramid<ImageType, AlphaType, PyramidType, 17, 7, mini::RGB<int32_t>, int16_t>(
                "lp", levels, false, img.upperLeft(), img.lowerRight(), img.accessor(), alpha.upperLeft(), alpha.accessor());
        FILE* o = fopen(argv[2], "wb");
        for (PyramidType* l : *lp)
            for (int c = 0; c < 3; ++c)
                for (int y = 0; y < l->height(); ++y)
                    for (int x = 0; x < l->width(); ++x) {
                        const int32_t v = (*l)(x, y)[c];
                        fwrite(&v, 4, 1, o);
                    }
        fclose(o);
        enblend::collapsePyramid<mini::RGB<int32_t>>(false, lp);
        long bad = 0;
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x)
                for (int c = 0; c < 3; ++c)
                    if ((*(*lp)[0])(x, y)[c] != ((int32_t)img(x, y)[c] << 7)) ++bad;
        for (PyramidType* l : *lp) delete l;
        delete lp;
        printf("dropin: %dx%d levels=%d collapse mismatches=%ld\n", w, h, levels, bad);
        return bad == 0 ? 0 : 1;
    } catch (const enblend_amd::error& e) {
        fprintf(stderr, "dropin: error %d: %s\n", e.code, e.what());
        return 3;
    }
}
