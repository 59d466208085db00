function [eigs, det1, err1] = gpu_multi_subsvd2(g, subIndex, minor)
%GPU_MULTI_SUBSVD2  As gpu_multi_subsvd, for index lists whose length is a
%   multiple of 32 (the reference's full-warp kernel; here the same
%   library call). Same signature, checks and outputs as
%   Matlab/gpu_multi_subsvd2.m:13-73.
eigs = -1; det1 = -1; err1 = -1;
[M, N] = size(g);
L = length(subIndex);
if M ~= N
    disp('dimensions of input for matrix don''t agree! ');
    return;
end
if mod(L, 32) ~= 0
    disp('length of subIndex must be divisible by 32! ');
    return;
end
if minor < 1 || floor(log2(minor)) - log2(minor) ~= 0
    disp('length of minor is not dyadic!');
    return;
end
if mod(L, minor) ~= 0
    disp('minor must divide into the length of the index! ');
    return;
end
if minor > 128
    disp('minor of length > 128 not supported! ');
    return;
end
[eigs, det1, err1] = pl_multi_subsvd_call(g, subIndex, minor);
end
