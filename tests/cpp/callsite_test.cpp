// The pyramid, blend and weight templates of include/enblend_amd/enblend.hpp
// driven the way enblend and enfuse drive them: the blend step of
// enblendMain (src/enblend.h:383-510) and the fusion loop of enfuseMain
// (src/enfuse.h:1246-1609) with the reference's template arguments
// (src/numerictraits.h:168-189), argument-object factories (srcImageRange,
// maskImage, destImage, ...) and a user exposure-weight function in the shape
// of doc/examples/enfuse/linear.cc.  Reads raw inputs from argv[2], writes
// the outputs to argv[3]:
//   enblend: w h levels, then white RGB16, white alpha, black RGB16, black
//            alpha, seam mask -> blended black RGB16, black alpha
//   enfuse8 / enfuse16: w h levels K, then K images and K alphas -> the fused
//            image, its alpha, and the exposure-weight table used
//   ops:     w h, then an RGB16 image and alpha -> the single operations'
//            results (see pyramid_ops)
// An optional argv[4] sets the blend colour space (ENBL_COLORSPACE_*), as
// --blend-colorspace sets BlendColorspace.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/enblend_amd/enblend.hpp"
#include "linear_weight.hpp"
#include "mini_image.hpp"

using namespace mini;

namespace {

template <class T>
bool read_vec(FILE* f, std::vector<T>& v)
{
    return fread(v.data(), sizeof(T), v.size(), f) == v.size();
}

template <class T>
void load_rgb(Image<RGB<T>>& im, const std::vector<T>& v)
{
    for (int y = 0; y < im.height(); ++y)
        for (int x = 0; x < im.width(); ++x) {
            const size_t i = ((size_t)y * im.width() + x) * 3;
            im(x, y) = RGB<T>(v[i], v[i + 1], v[i + 2]);
        }
}

template <class T>
void load_plane(Image<T>& im, const std::vector<T>& v)
{
    for (int y = 0; y < im.height(); ++y)
        for (int x = 0; x < im.width(); ++x) im(x, y) = v[(size_t)y * im.width() + x];
}

template <class T>
void write_rgb(FILE* f, const Image<RGB<T>>& im)
{
    for (int y = 0; y < im.height(); ++y)
        for (int x = 0; x < im.width(); ++x)
            for (int c = 0; c < 3; ++c) {
                const T v = im(x, y)[c];
                fwrite(&v, sizeof(T), 1, f);
            }
}

template <class T>
void write_plane(FILE* f, const Image<T>& im)
{
    for (int y = 0; y < im.height(); ++y)
        for (int x = 0; x < im.width(); ++x) fwrite(&im(x, y), sizeof(T), 1, f);
}

template <class P>
void delete_pyramid(std::vector<P*>* p)
{
    for (P* l : *p) delete l;
    delete p;
}

// One iteration of enblendMain's loop from the mask on (src/enblend.h:383-510),
// ROI = the whole canvas, 16-bit RGB images.
int enblend_step(FILE* in, FILE* out)
{
    typedef Image<RGB<uint16_t>> ImageType;
    typedef Image<uint8_t> AlphaType;
    typedef Image<uint8_t> MaskType;
    typedef Image<RGB<int32_t>> ImagePyramidType;
    typedef Image<int32_t> MaskPyramidType;
    int32_t hdr[3];
    if (fread(hdr, 4, 3, in) != 3) return 2;
    const int w = hdr[0], h = hdr[1];
    const unsigned numLevels = (unsigned)hdr[2];
    std::vector<uint16_t> pw((size_t)w * h * 3), pb((size_t)w * h * 3);
    std::vector<uint8_t> aw((size_t)w * h), ab((size_t)w * h), mk((size_t)w * h);
    if (!read_vec(in, pw) || !read_vec(in, aw) || !read_vec(in, pb) || !read_vec(in, ab) || !read_vec(in, mk)) return 2;
    ImageType white(w, h), black(w, h);
    AlphaType whiteAlpha(w, h), blackAlpha(w, h);
    MaskType mask(w, h);
    load_rgb(white, pw);
    load_rgb(black, pb);
    load_plane(whiteAlpha, aw);
    load_plane(blackAlpha, ab);
    load_plane(mask, mk);
    const bool wraparoundForBlend = false;

    // Gaussian pyramid of the seam mask
    std::vector<MaskPyramidType*>* maskGP =
        enblend::gaussianPyramid<MaskType, MaskPyramidType, 9, 15, int32_t>(numLevels, wraparoundForBlend,
                                                                            srcImageRange(mask));
    // Laplacian pyramids of white and black under their own alpha
    std::vector<ImagePyramidType*>* whiteLP =
        enblend::laplacianPyramid<ImageType, AlphaType, ImagePyramidType, 17, 7, RGB<int32_t>, int16_t>(
            "whiteGP", numLevels, wraparoundForBlend, srcImageRange(white), maskImage(whiteAlpha));
    std::vector<ImagePyramidType*>* blackLP =
        enblend::laplacianPyramid<ImageType, AlphaType, ImagePyramidType, 17, 7, RGB<int32_t>, int16_t>(
            "blackGP", numLevels, wraparoundForBlend, srcImageRange(black), maskImage(blackAlpha));
    // black alpha becomes the union of both alphas
    initImageIf(destImageRange(blackAlpha), maskImage(whiteAlpha), (uint8_t)255);
    // blend with white = the mask maximum converted to the mask pyramid
    enblend::blend(maskGP, whiteLP, blackLP, (int32_t)(255 << 15));
    delete_pyramid(maskGP);
    delete_pyramid(whiteLP);
    enblend::collapsePyramid<RGB<int32_t>>(wraparoundForBlend, blackLP);
    enblend::copyFromPyramidImageIf<ImagePyramidType, MaskType, ImageType, 17, 7>(
        srcImageRange(*((*blackLP)[0])), maskImage(blackAlpha), destImage(black));
    delete_pyramid(blackLP);
    write_rgb(out, black);
    write_plane(out, blackAlpha);
    return 0;
}

// enfuseMain's mask generation and fusion (src/enfuse.h:1246-1385,
// :1497-1609), soft masks, for 8- or 16-bit RGB images.
template <class T, class PT, class MPT, int IB, int FB, int MIB, int MFB>
int enfuse(FILE* in, FILE* out)
{
    typedef Image<RGB<T>> ImageType;
    typedef Image<uint8_t> AlphaType;
    typedef Image<float> MaskType;
    typedef Image<RGB<PT>> ImagePyramidType;
    typedef Image<MPT> MaskPyramidType;
    int32_t hdr[4];
    if (fread(hdr, 4, 4, in) != 4) return 2;
    const int w = hdr[0], h = hdr[1], K = hdr[3];
    const unsigned numLevels = (unsigned)hdr[2];

    // the user weight function, made as make_weight_function makes a plugin
    static LinearWeight linear;
    const std::vector<std::string> args;
    linear.initialize(0.5, 0.2, args.begin(), args.end());
    enblend_amd::EnfuseOptions& opt = enblend_amd::enfuse_options();
    opt.WExposure = 1.0;
    opt.WSaturation = 0.2;
    opt.WContrast = 0.3;
    opt.ContrastWindowSize = 5;
    opt.ExposureWeightFunction = &linear;

    std::vector<ImageType*> images;
    std::vector<AlphaType*> alphas;
    std::vector<MaskType*> masks;
    MaskType normImage(w, h);
    AlphaType outputAlpha(w, h);
    for (int m = 0; m < K; ++m) {
        std::vector<T> px((size_t)w * h * 3);
        std::vector<uint8_t> al((size_t)w * h);
        if (!read_vec(in, px)) return 2;
        images.push_back(new ImageType(w, h));
        load_rgb(*images.back(), px);
        (void)al;
    }
    for (int m = 0; m < K; ++m) {
        std::vector<uint8_t> al((size_t)w * h);
        if (!read_vec(in, al)) return 2;
        alphas.push_back(new AlphaType(w, h));
        load_plane(*alphas.back(), al);
    }
    for (int m = 0; m < K; ++m) {
        MaskType* mask = new MaskType(w, h);
        enblend::enfuseMask<ImageType, AlphaType, MaskType>(srcImageRange(*images[m]), maskImage(*alphas[m]),
                                                            destImage(*mask));
        // output alpha: union of the input alphas; norm += mask
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                if ((*alphas[m])(x, y)) outputAlpha(x, y) = (*alphas[m])(x, y);
                normImage(x, y) = (*mask)(x, y) + normImage(x, y);
            }
        masks.push_back(mask);
    }
    const uint8_t maxMaskPixelType = 255;
    std::vector<ImagePyramidType*>* resultLP = nullptr;
    for (int m = 0; m < K; ++m) {
        std::vector<ImagePyramidType*>* imageLP =
            enblend::laplacianPyramid<ImageType, AlphaType, ImagePyramidType, IB, FB, RGB<int32_t>, int16_t>(
                "imageGP", numLevels, false, srcImageRange(*images[m]), maskImage(*alphas[m]));
        // normalise the mask coefficients to the mask pixel range
        MaskType& mask = *masks[m];
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x)
                mask(x, y) = normImage(x, y) > 0.0f ? (float)maxMaskPixelType * mask(x, y) / normImage(x, y)
                                                    : (float)(maxMaskPixelType / K);
        std::vector<MaskPyramidType*>* maskGP =
            enblend::gaussianPyramid<MaskType, AlphaType, MaskPyramidType, MIB, MFB, int32_t, int16_t>(
                numLevels, false, srcImageRange(mask), maskImage(outputAlpha));
        const MPT maxMaskPyramidPixelValue = (MPT)(maxMaskPixelType << MFB);
        const enblend::ImageMaskMultiplyFunctor<MPT> multiply(maxMaskPyramidPixelValue);
        for (size_t l = 0; l < maskGP->size(); ++l) {
            ImagePyramidType& lp = *(*imageLP)[l];
            const MaskPyramidType& mp = *(*maskGP)[l];
            for (int y = 0; y < lp.height(); ++y)
                for (int x = 0; x < lp.width(); ++x) lp(x, y) = multiply(lp(x, y), mp(x, y));
        }
        delete_pyramid(maskGP);
        if (resultLP) {
            // resultLP += imageLP, stored with the pyramid type's wrap
            for (size_t l = 0; l < imageLP->size(); ++l) {
                ImagePyramidType& r = *(*resultLP)[l];
                const ImagePyramidType& a = *(*imageLP)[l];
                for (int y = 0; y < r.height(); ++y)
                    for (int x = 0; x < r.width(); ++x)
                        for (int c = 0; c < 3; ++c)
                            r(x, y)[c] = (PT)(uint32_t)((uint32_t)(int32_t)a(x, y)[c] + (uint32_t)(int32_t)r(x, y)[c]);
            }
            delete_pyramid(imageLP);
        } else {
            resultLP = imageLP;
        }
    }
    enblend::collapsePyramid<RGB<int32_t>>(false, resultLP);
    ImageType output(w, h);
    enblend::copyFromPyramidImageIf<ImagePyramidType, AlphaType, ImageType, IB, FB>(
        srcImageRange(*((*resultLP)[0])), maskImage(outputAlpha), destImage(output));
    delete_pyramid(resultLP);
    write_rgb(out, output);
    write_plane(out, outputAlpha);
    const std::vector<double> table = enblend_amd::exposure_weight_table(&linear, sizeof(T) == 1 ? 256 : 65536);
    fwrite(table.data(), sizeof(double), table.size(), out);
    for (auto* p : images) delete p;
    for (auto* p : alphas) delete p;
    for (auto* p : masks) delete p;
    return 0;
}

// The single operations through their argument-object forms, 16-bit RGB:
// copyToPyramidImage, reduce with and without alpha, expand subtracting and
// adding.  Writes level 0, the reduced level and its alpha, the reduced level
// without alpha, level 0 minus and plus the expansion of the reduced level.
int pyramid_ops(FILE* in, FILE* out)
{
    typedef Image<RGB<uint16_t>> ImageType;
    typedef Image<uint8_t> AlphaType;
    typedef Image<RGB<int32_t>> ImagePyramidType;
    int32_t hdr[2];
    if (fread(hdr, 4, 2, in) != 2) return 2;
    const int w = hdr[0], h = hdr[1], cw = (w + 1) / 2, ch = (h + 1) / 2;
    std::vector<uint16_t> px((size_t)w * h * 3);
    std::vector<uint8_t> al((size_t)w * h);
    if (!read_vec(in, px) || !read_vec(in, al)) return 2;
    ImageType img(w, h);
    AlphaType alpha(w, h);
    load_rgb(img, px);
    load_plane(alpha, al);
    ImagePyramidType l0(w, h), r1(cw, ch), r2(cw, ch), sub(w, h), add(w, h);
    AlphaType a1(cw, ch);
    enblend::copyToPyramidImage<ImageType, ImagePyramidType, 17, 7>(srcImageRange(img), destImage(l0));
    enblend::reduce<RGB<int32_t>, int16_t>(false, srcImageRange(l0), maskImage(alpha), destImageRange(r1),
                                           destImageRange(a1));
    enblend::reduce<RGB<int32_t>>(false, srcImageRange(l0), destImageRange(r2));
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) sub(x, y) = add(x, y) = l0(x, y);
    enblend::expand<RGB<int32_t>>(false, false, srcImageRange(r1), destImageRange(sub));
    enblend::expand<RGB<int32_t>>(true, false, srcImageRange(r1), destImageRange(add));
    write_rgb(out, l0);
    write_rgb(out, r1);
    write_plane(out, a1);
    write_rgb(out, r2);
    write_rgb(out, sub);
    write_rgb(out, add);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: callsite_test enblend|enfuse8|enfuse16|ops in.bin out.bin [colorspace]\n");
        return 2;
    }
    FILE* in = fopen(argv[2], "rb");
    FILE* out = fopen(argv[3], "wb");
    if (!in || !out) return 2;
    int rc = 2;
    if (argc > 4) enblend_amd::blend_colorspace() = atoi(argv[4]);
    try {
        if (!strcmp(argv[1], "enblend")) rc = enblend_step(in, out);
        else if (!strcmp(argv[1], "enfuse8")) rc = enfuse<uint8_t, int16_t, int16_t, 9, 7, 9, 7>(in, out);
        else if (!strcmp(argv[1], "enfuse16")) rc = enfuse<uint16_t, int32_t, int32_t, 17, 7, 9, 15>(in, out);
        else if (!strcmp(argv[1], "ops")) rc = pyramid_ops(in, out);
    } catch (const enblend_amd::error& e) {
        fprintf(stderr, "callsite_test: enblend_amd error %d: %s\n", e.code, e.what());
        rc = 3;
    }
    fclose(in);
    fclose(out);
    return rc;
}
