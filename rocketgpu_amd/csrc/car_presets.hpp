// The stock car bodies as CarConfig records: the game's published hitboxes (with Bullet's 2.5 uu margin on the
// collision box), wheel layouts, suspension rest lengths, steering curves and spawn heights, as the reference
// defines them in OctaneConfig() ... BatmobileConfig() (include/rocketgpu/rocketgpu.h:1098-1110,
// src/rocketgpu.cpp:417-520).  These numbers are the interface's data: a car built from them must be the same
// bytes as the reference's.  Fields a body does not set keep the Octane defaults of CarConfig (rg_types.h): the
// steering curves of the Merc, and the published extent and steering curve of the Psyclops.
#pragma once

#include "rocketgpu_amd/rg_types.h"

namespace rg {

enum CarPreset : int {
    kPresetOctane = 0,
    kPresetDominus,
    kPresetPlank,
    kPresetBreakout,
    kPresetHybrid,
    kPresetMerc,
    kPresetPsyclops,
    kPresetBatmobile,
    kPresetCount
};

namespace presets {

struct Wheels {
    float radius, rest;
    Vec3 offset;
};

struct Body {
    Vec3 published;       // the published hitbox extent; the collision box is 2.5 uu larger on each axis
    Vec3 offset;          // hitbox offset
    Wheels front, back;
    bool three_wheels;
    bool sets_published;  // false: CarConfig's default published extent stays
    const float* steer;   // 6 values, or null for the default curve
    const float* slide;   // 2 values, or null for the default curve
    float spawn_z;
    uint16_t body;
};

constexpr float kDominusSteer[6] = {29.923809051513672f, 18.018766403198242f, 10.329229354858398f,
                                    6.019903182983398f, 4.85035514831543f, 1.9751166105270386f};
constexpr float kDominusSlide[2] = {22.08690071105957f, 7.173928260803223f};
constexpr float kPlankSteer[6] = {30.342710494995117f, 18.1981258392334f, 10.394840240478516f,
                                  6.043498516082764f, 4.865903377532959f, 1.9777874946594238f};
constexpr float kPlankSlide[2] = {22.341325759887695f, 7.206938743591309f};
constexpr float kBreakoutSteer[6] = {30.55976104736328f, 18.289745330810547f, 10.42794418334961f,
                                     6.055309772491455f, 4.873669147491455f, 1.9791138172149658f};
constexpr float kBreakoutSlide[2] = {22.472003936767578f, 7.2235002517700195f};
constexpr float kHybridSteer[6] = {30.580455780029297f, 18.29843521118164f, 10.431071281433105f,
                                   6.056422233581543f, 4.874399185180664f, 1.9792383909225464f};
constexpr float kHybridSlide[2] = {22.484424591064453f, 7.225060939788818f};
constexpr float kPsyclopsSlide[2] = {24.427291870117188f, 7.443766117095947f};
constexpr float kBatmobileSteer[6] = {30.342700958251953f, 18.1981201171875f, 10.394838333129883f,
                                      6.0434980392456055f, 4.865902900695801f, 1.9777874946594238f};
constexpr float kBatmobileSlide[2] = {22.341318130493164f, 7.206938743591309f};

inline Body BodyOf(int preset) {
    switch (preset) {
    case kPresetDominus:
        return {Vec3(127.92678f, 83.27995f, 31.3f), Vec3(9.0f, 0.0f, 15.75f),
                {12.00f, 33.95f, Vec3(50.30f, 31.10f, 15.75f)}, {13.50f, 33.85f, Vec3(-34.75f, 33.00f, 15.75f)},
                false, true, kDominusSteer, kDominusSlide, 36.01f, 132};
    case kPresetPlank:
        return {Vec3(128.81978f, 84.670364f, 29.394402f), Vec3(9.008572f, 0.0f, 12.0942f),
                {12.50f, 31.9242f, Vec3(49.97f, 27.799904f, 12.0942f)},
                {17.00f, 27.9242f, Vec3(-35.43f, 20.280005f, 12.0942f)}, false, true, kPlankSteer, kPlankSlide,
                38.84f, 150};
    case kPresetBreakout:
        return {Vec3(131.49236f, 80.521f, 30.3f), Vec3(12.5f, 0.0f, 11.75f),
                {13.50f, 29.7f, Vec3(51.50f, 26.67f, 11.75f)}, {15.00f, 29.66f, Vec3(-35.75f, 35.00f, 11.75f)},
                false, true, kBreakoutSteer, kBreakoutSlide, 36.00f, 64};
    case kPresetHybrid:
        return {Vec3(127.01919f, 82.18787f, 34.159073f), Vec3(13.87566f, 0.0f, 20.754988f),
                {12.50f, 38.75499f, Vec3(51.25f, 25.90f, 20.754988f)},
                {15.00f, 37.054989f, Vec3(-34.00f, 29.50f, 20.754988f)}, false, true, kHybridSteer, kHybridSlide,
                37.53f, 56};
    case kPresetMerc:
        return {Vec3(120.72023f, 76.71031f, 41.659073f), Vec3(11.37566f, 0.0f, 21.504988f),
                {15.00f, 39.50499f, Vec3(51.25f, 25.90f, 21.504988f)},
                {15.00f, 39.104988f, Vec3(-33.75f, 29.50f, 21.504988f)}, false, true, nullptr, nullptr, 35.05f,
                219};
    case kPresetPsyclops:
        return {Vec3(118.00738f, 84.19941f, 36.159073f), Vec3(13.87566f, 0.0f, 15.00499f),
                {12.50f, 33.00499f, Vec3(51.25f, 5.000f, 15.00499f)},
                {15.00f, 31.304989f, Vec3(-34.25f, 29.50f, 15.00499f)}, true, false, nullptr, kPsyclopsSlide,
                36.00f, 214};
    case kPresetBatmobile:
        return {Vec3(128.81978f, 84.670364f, 29.394402f), Vec3(9.008572f, 0.0f, 12.0942f),
                {12.50f, 31.9242f, Vec3(49.97f, 27.80f, 12.0942f)}, {17.00f, 27.9242f, Vec3(-35.43f, 20.28f, 12.0942f)},
                false, true, kBatmobileSteer, kBatmobileSlide, 36.00f, 41};
    default:
        return {};
    }
}

}  // namespace presets

// The CarConfig of a stock body; Octane (CarConfig's defaults) for kPresetOctane.
inline CarConfig CarPresetConfig(int preset) {
    CarConfig c{};
    if (preset == kPresetOctane) return c;
    const presets::Body b = presets::BodyOf(preset);
    c.hitbox_size = Vec3(b.published.x + 2.5f, b.published.y + 2.5f, b.published.z + 2.5f);
    if (b.sets_published) c.hitbox_published = b.published;
    c.hitbox_offset = b.offset;
    c.front_wheels_offset = b.front.offset;
    c.back_wheels_offset = b.back.offset;
    c.front_wheel_radius = b.front.radius;
    c.back_wheel_radius = b.back.radius;
    c.front_suspension_rest = b.front.rest;
    c.back_suspension_rest = b.back.rest;
    c.three_wheels = b.three_wheels ? 1 : 0;
    if (b.steer) {
        for (int i = 0; i < 6; ++i) c.steer_angle_curve_deg[i] = b.steer[i];
    }
    if (b.slide) {
        for (int i = 0; i < 2; ++i) c.powerslide_steer_angle_curve_deg[i] = b.slide[i];
    }
    c.spawn_z = b.spawn_z;
    c.body = b.body;
    return c;
}

}  // namespace rg
