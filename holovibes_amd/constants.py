"""Constants of the C-ABI (include/holo_hip.h), mirroring the reference enums."""

# camera::PixelDepth (Camera/include/frame_desc.hh:13-24)
HV_DEPTH_8, HV_DEPTH_16, HV_DEPTH_32F = 1, 2, 4
# holovibes::SpaceTransformation (Backend/includes/enum/enum_space_transformation.hh:17-22)
HV_SPACE_NONE, HV_SPACE_FRESNELTR, HV_SPACE_ANGULARSP = 0, 1, 2
# holovibes::TimeTransformation (Backend/includes/enum/enum_time_transformation.hh:16-22)
HV_TIME_NONE, HV_TIME_STFT, HV_TIME_PCA = 0, 1, 2
# holovibes::ImgType (Backend/includes/enum/enum_img_type.hh:15-25)
HV_IMG_MODULUS, HV_IMG_SQUARED_MODULUS, HV_IMG_ARGUMENT = 0, 1, 2
HV_IMG_MOMENTS_0, HV_IMG_MOMENTS_1, HV_IMG_MOMENTS_2 = 5, 6, 7
# lens arithmetic
HV_LENS_FP64, HV_LENS_REFERENCE_FP32 = 0, 1
# PCA covariance arithmetic
HV_PCA_FP32, HV_PCA_BF16_SPLIT, HV_PCA_FP16_SPLIT, HV_PCA_BF16, HV_PCA_FP16 = 0, 1, 2, 3, 4
# contrast policy
HV_CONTRAST_REFERENCE, HV_CONTRAST_PERCENTILE, HV_CONTRAST_MINMAX, HV_CONTRAST_MANUAL = 0, 1, 2, 3
